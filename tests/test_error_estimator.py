"""The error report's estimator on the CPU (the host replica in _report.py;
the GPU tests check that the library computes the same number).  Synthetic
fields of 256^2 targets -- random phases (energy spread evenly), one focus,
and 64 bright spots on a dark background as an optimised hologram makes --
carry synthetic errors: a part proportional to the field (as errors that add
coherently at a focus do) and a part of even absolute size (as incoherent
rounding does).  The report (half its targets drawn by energy, half spread
evenly, importance-weighted) stays within a factor 1.4 of the whole field's
relative L2 error in every case; a stride subset's own
relative error misses it by a wide margin on the spot field, and the stride
subset's error over the whole field's norm misses it at the focus."""
import numpy as np
import pytest

from _report import pick, rel_l2, report_of, tmap

N = 256


def fields():
    g = np.random.default_rng(0)
    yy, xx = np.mgrid[0:N, 0:N]
    rnd = g.standard_normal(N * N) + 1j * g.standard_normal(N * N)
    focus = (np.exp(-((xx - 77.3) ** 2 + (yy - 140.6) ** 2) / (2 * 4.0 ** 2))
             * np.exp(1j * 0.01 * xx)).ravel() + 1e-3 * rnd
    spots = np.zeros((N, N), complex)
    for cx, cy in g.uniform(8, N - 8, (64, 2)):
        spots += np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * 1.5 ** 2) + 1j * g.uniform(0, 6.3))
    spots = spots.ravel() + 3e-2 * rnd
    return {"random": rnd, "focus": focus, "spots": spots}


@pytest.mark.parametrize("name", ["random", "focus", "spots"])
@pytest.mark.parametrize("coh,inc", [(1e-7, 0.0), (0.0, 1e-7), (1e-7, 1e-7), (3e-8, 2e-7)])
def test_energy_drawn_report_follows_whole_error(name, coh, inc):
    y = fields()[name]
    g = np.random.default_rng(1)
    rms = np.sqrt(np.mean(np.abs(y) ** 2))
    noise = (g.standard_normal(y.size) + 1j * g.standard_normal(y.size)) / np.sqrt(2)
    r = y * (1 + coh * np.exp(1j * 0.3)) + inc * rms * noise
    whole = rel_l2(r, y)
    pk = pick(r)
    assert pk[1]
    rep = report_of(r, pk, y[pk[0]])
    assert whole / 1.4 <= rep <= 1.4 * whole, (rep, whole)


def test_stride_estimates_fail_on_concentrated_fields():
    f = fields()
    g = np.random.default_rng(2)
    st = tmap(N * N)
    # spots with even absolute error: the stride subset's own norm overstates
    y = f["spots"]
    rms = np.sqrt(np.mean(np.abs(y) ** 2))
    r = y + 1e-7 * rms * (g.standard_normal(y.size) + 1j * g.standard_normal(y.size))
    assert rel_l2(r[st], y[st]) > 2 * rel_l2(r, y)
    # focus with error proportional to the field: the subset's error over the
    # whole norm understates
    y = f["focus"]
    r = y * (1 + 1e-7j)
    whole = rel_l2(r, y)
    est = np.sqrt(np.sum(np.abs(r[st] - y[st]) ** 2) * (y.size / st.size) / np.sum(np.abs(r) ** 2))
    assert est < whole / 3


def test_zero_and_tiny_fields():
    """No energy: stride targets; one bright target: every draw takes it."""
    z = np.zeros(1000, complex)
    sub, weighted, ne = pick(z)
    assert not weighted and ne == 0 and np.array_equal(sub, tmap(1000))
    one = np.zeros(1000, complex)
    one[417] = 2.0
    sub, weighted, ne = pick(one)
    assert weighted and ne == 128 and np.all(sub[:ne] == 417)
    assert np.array_equal(sub[ne:], tmap(1000, 128))
    # M <= 256: every target once, and the report is the exact value
    y = np.exp(1j * np.arange(200.0))
    r = y * (1 + 1e-7) + 1e-7
    pk = pick(r)
    assert pk[2] == 0 and np.array_equal(pk[0], np.arange(200))
    assert abs(report_of(r, pk, y) - np.linalg.norm(r - y) / np.linalg.norm(r)) < 1e-12
