"""The committed fixtures in tests/golden/ against closed forms (CPU).

matlab_projector_*: the input set of matlab/projector.m:4-75 at 32x32 (one
unit point source at (0, 0, 0.5 m), a non-centred 7 um projector lattice at
z = 0, x slow / y fast, holofunctions-style `flatten_positions` AoS layout).
With a single source the CLI's outputs have closed forms: after
normalisation |y_m| = d_min / d_m, and the phase is -+ K d_m (mod 2 pi) for
Backwards / Forwards, K = (2 pi) / lambda (cuda/macros.h:29-31).
"""
import os

import mpmath as mp
import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LAM_REF = 0.65e-6


def dat(name):
    return np.fromfile(os.path.join(GOLDEN, name), "<f8")


def test_matlab_projector_layout():
    v = dat("matlab_projector_v0.dat").reshape(-1, 3)
    w = dat("matlab_projector_w0.dat").reshape(-1, 3)
    assert v.shape == (1024, 3) and w.shape == (1024, 3)
    i, j = np.divmod(np.arange(1024), 32)  # y + (x-1) * Ny: x slow, y fast
    assert np.array_equal(v[:, 0], i * 7e-6) and np.array_equal(v[:, 1], j * 7e-6)
    assert np.all(v[:, 2] == 0.0) and np.all(w[:, 2] == 0.5)
    assert np.array_equal(w[:, :2], v[:, :2])
    assert dat("matlab_projector_u0.dat").tolist() == [0.0, 0.0, 0.5]
    assert dat("matlab_projector_x0_amp.dat").tolist() == [1.0]
    assert dat("matlab_projector_x0_phase.dat").tolist() == [0.0]


@pytest.mark.parametrize("name, sign", [("y0", -1), ("z0_0", +1)])
def test_matlab_projector_outputs_closed_form(name, sign):
    v = dat("matlab_projector_v0.dat").reshape(-1, 3)
    amp = dat(f"matlab_projector_{name}_amp.dat")
    phase = dat(f"matlab_projector_{name}_phase.dat")
    K = mp.mpf(2) * mp.pi / mp.mpf(LAM_REF)
    d = [mp.sqrt(mp.mpf(x) ** 2 + mp.mpf(y) ** 2 + mp.mpf(0.5) ** 2) for x, y, _ in v]
    dmin = min(d)
    for m in range(0, 1024, 7):
        assert amp[m] == pytest.approx(float(dmin / d[m]), rel=1e-15)
        expect = float(mp.fmod(sign * K * d[m], 2 * mp.pi))
        # one FP64 rounding of the ~4.8e6 rad argument: ~5e-10 rad
        assert abs(np.angle(np.exp(1j * (phase[m] - expect)))) < 2e-9
    assert np.all(phase > -np.pi) and np.all(phase <= np.pi)
    assert amp.max() == 1.0
