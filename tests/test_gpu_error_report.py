"""The error report of every result (err_rel_l2 of holo.h).

An FP32-split call also evaluates the FP64 formulation at count = min(M, 256)
targets drawn by the energy of its result r: m_i is the first target at which
the running sum of |r|^2 exceeds (i + 1/2) / count of the total, and the
report is sqrt(mean_i |r[m_i] - r64_i|^2 / |r[m_i]|^2), an estimate of the
relative L2 error of the whole result (its expectation under a random offset
of the strata).  A result without positive finite energy falls back to the
stride targets floor(i M / count) and their relative L2 error.  Each test
replays the choice on the host, evaluates the oracle at the chosen targets
and requires the two reports to agree to 1e-3 of the value (the library's
FP64 values are within 1e-12 of the oracle's); FP64 calls report 0.  Host
and device pointers both carry it.  A focused field, and a field whose energy
lies away from any stride subset, check that the report follows the whole
result's error.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LAM = 632.8e-9


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.nodeid}", flush=True)
    sys.stdout.flush()
    yield


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def tmap(M, count=256):
    count = min(M, count)
    return (np.arange(count, dtype=np.int64) * M) // count


def pick(r, count=256):
    """The report's targets for result r (all M values), as the library picks them."""
    r = np.asarray(r)
    M = len(r)
    count = min(M, count)
    e = np.abs(r) ** 2
    C = np.cumsum(e)
    total = C[-1]
    if not (total > 0 and np.isfinite(total)):
        return tmap(M, count), False
    c = (np.arange(count) + 0.5) / count * total
    m = np.searchsorted(C, c, side="right")
    last = int(np.flatnonzero(e > 0)[-1])
    return np.where(m < M, m, last).astype(np.int64), True


def report_of(r, sub, ref, weighted=True):
    """The report's formula on the host: r (all M values), ref = r64 at sub."""
    r = np.asarray(r)
    a, b = r[sub], np.asarray(ref)
    if weighted:
        return float(np.sqrt(np.mean(np.abs(a - b) ** 2 / np.abs(a) ** 2)))
    return rel_l2(a, b)


def agree(reported, host):
    assert 0.0 < host <= 1e-6
    assert abs(reported - host) <= 1e-3 * host + 1e-15, (reported, host)


def to_dev(*arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in arrays]


@pytest.mark.parametrize("geometry", ["lattice", "cloud", "padded"])
@pytest.mark.parametrize("M", [2000, 97])
def test_transform_full_and_plan(pkg, ctx, oracle, workloads, geometry, M):
    g = np.random.default_rng(M)
    if geometry == "lattice":
        u = workloads.grid(96, 64, 10e-6, 0.0)
    elif geometry == "padded":
        u = workloads.reference_plane(5000, 1.0e-3, 0.0, 16 / 9)
    else:
        u = np.zeros((6000, 3))
        u[:, :2] = (g.random((6000, 2)) - 0.5) * 2e-3
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    v = np.zeros((M, 3))
    v[:, :2] = (g.random((M, 2)) - 0.5) * 3e-3
    v[:, 2] = 0.3
    # transform_full, host pointers
    y = ctx.transform_full(amp, phase, u, v, -1, LAM, pkg.FP32_SPLIT, report_error=True)
    sub, _ = pick(y)
    ref = oracle.transform_full(-1, LAM, amp, phase, u, v, targets=sub)
    agree(ctx.last_rel_l2, report_of(y, sub, ref))
    # plan, device pointers (the report stays on the device)
    da, dp, du, dv = to_dev(amp, phase, u, v)
    plan = ctx.plan(du)
    yd = plan.transform(da, dp, dv, -1, LAM, pkg.FP32_SPLIT, report_error=True)
    e = ctx.last_rel_l2
    assert hasattr(e, "device") and e.device.type == "cuda"
    yd = yd.cpu().numpy()
    subd, _ = pick(yd)
    refd = oracle.transform_full(-1, LAM, amp, phase, u, v, targets=subd)
    agree(float(e.item()), report_of(yd, subd, refd))
    # FP64 reports 0
    ctx.transform_full(amp, phase, u, v, -1, LAM, pkg.FP64, report_error=True)
    assert ctx.last_rel_l2 == 0.0
    plan.transform(da, dp, dv, -1, LAM, pkg.FP64, report_error=True)
    assert float(ctx.last_rel_l2.item()) == 0.0


def test_plan_complex_weights_and_obliquity(pkg, ctx, oracle, workloads):
    u = workloads.grid(64, 64, 10e-6, 0.0)
    g = np.random.default_rng(1)
    w = g.standard_normal(len(u)) + 1j * g.standard_normal(len(u))
    v = workloads.grid(40, 30, 7.56e-6, 0.25)
    plan = ctx.plan(u)
    y = plan.transform_complex(w, v, +1, LAM, pkg.FP32_SPLIT, obliquity=True, report_error=True)
    sub, _ = pick(y)
    ref = oracle.transform_cartesian(+1, LAM, w, u, v[sub], obliquity=True)
    agree(ctx.last_rel_l2, report_of(y, sub, ref))


@pytest.mark.parametrize("shared", [False, True])
@pytest.mark.parametrize("lattice", [False, True])
def test_monte_carlo(pkg, ctx, oracle, workloads, shared, lattice):
    """The FP64 estimator with the same indices at the stride targets."""
    g = np.random.default_rng(2)
    if lattice:
        u = workloads.grid(100, 80, 10e-6, 0.0)
    else:
        u = np.zeros((8000, 3))
        u[:, :2] = (g.random((8000, 2)) - 0.5) * 2e-3
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    v = workloads.grid(50, 41, 7.56e-6, 0.3)
    K = 333
    y = ctx.transform_mc(amp, phase, u, v, K, 9, +1, LAM, pkg.FP32_SPLIT, shared=shared,
                         report_error=True)
    sub, _ = pick(y)
    ref = oracle.transform_mc(+1, LAM, amp, phase, u, v, K, 9, targets=sub, shared=shared)
    agree(ctx.last_rel_l2, report_of(y, sub, ref))
    ctx.transform_mc(amp, phase, u, v, K, 9, +1, LAM, pkg.FP64, shared=shared,
                     report_error=True)
    assert ctx.last_rel_l2 == 0.0


def test_adjoint(pkg, ctx, oracle, workloads):
    """The report covers b_n = sum_m g_m conj(G_mn); b is recovered from the
    gradients as e^{i theta} (grad_amp + i grad_theta / amp)."""
    u = workloads.grid(64, 64, 7.56e-6, 0.0)
    v = workloads.grid(32, 32, 10e-6, 0.35)
    g = np.random.default_rng(3)
    amp, theta = 0.5 + g.random(len(u)), g.random(len(u)) * 2 * np.pi
    gz = g.standard_normal(len(v)) + 1j * g.standard_normal(len(v))
    plan_v = ctx.plan(v)
    ga, gt = plan_v.adjoint(gz, amp, theta, u, +1, LAM, pkg.FP32_SPLIT, report_error=True)
    b = np.exp(1j * theta) * (ga + 1j * gt / amp)
    sub, _ = pick(b)
    rb = oracle.transform_cartesian(-1, LAM, gz, v, u[sub])
    agree(ctx.last_rel_l2, report_of(b, sub, rb))


def test_focused_field_report_follows_whole_result(pkg, ctx, oracle, workloads):
    """Sources phased to focus near target 1, every 8th target (a stride
    subset of 2048) 2 mm to the side in the dark: the report stays within a
    factor 2 of the whole result's relative L2 error against the oracle and
    samples the bright targets, while the stride subset's own relative error
    would be far higher."""
    u = workloads.grid(96, 64, 10e-6, 0.0)
    M = 2048
    v = np.zeros((M, 3))
    g = np.random.default_rng(7)
    v[:, :2] = (g.random((M, 2)) - 0.5) * 60e-6       # around the focus
    v[::8, 0] += 2e-3                                  # in the dark
    v[:, 2] = 0.3
    k = 2 * np.pi / LAM
    phase = np.mod(-k * np.linalg.norm(u - v[1], axis=1), 2 * np.pi)
    amp = np.ones(len(u))
    y = ctx.transform_full(amp, phase, u, v, +1, LAM, pkg.FP32_SPLIT, report_error=True)
    rep = ctx.last_rel_l2
    full = oracle.transform_full(+1, LAM, amp, phase, u, v)
    sub, weighted = pick(y)
    assert weighted and np.mean(sub % 8 == 0) < 0.05
    agree(rep, report_of(y, sub, full[sub]))
    whole = rel_l2(y, full)
    assert 0.5 * whole <= rep <= 2.0 * whole, (rep, whole)
    stride = tmap(M)
    assert rel_l2(y[stride], full[stride]) > 5.0 * whole


def test_zero_result_reports_stride_error(pkg, ctx, oracle):
    """A result with no energy (all amplitudes 0) takes the stride targets:
    0 against an FP64 result of 0."""
    g = np.random.default_rng(8)
    u = np.zeros((500, 3))
    u[:, :2] = (g.random((500, 2)) - 0.5) * 1e-3
    v = np.zeros((300, 3))
    v[:, :2] = (g.random((300, 2)) - 0.5) * 1e-3
    v[:, 2] = 0.2
    y = ctx.transform_full(np.zeros(500), g.random(500), u, v, +1, LAM, pkg.FP32_SPLIT,
                           report_error=True)
    assert np.all(y == 0) and ctx.last_rel_l2 == 0.0


def test_cpp_dropin_reports(pkg, ctx):
    """The C ABI directly: a NULL report pointer is allowed; FP32-split with
    zero sources reports 0."""
    import ctypes
    L = pkg.lib()
    y = np.empty(4)
    v = np.array([0.0, 0.0, 0.1, 1e-3, 0.0, 0.1])
    e = ctypes.c_double(-1.0)
    assert L.holo_transform_full(ctx.handle, pkg.FP32_SPLIT, 1, LAM, None, None, None, 0,
                                 v.ctypes.data, 2, y.ctypes.data, ctypes.addressof(e), 0) == 0
    assert e.value == 0.0 and np.all(y == 0.0)
    one = np.ones(1)
    u = np.zeros(3)
    assert L.holo_transform_full(ctx.handle, pkg.FP32_SPLIT, 1, LAM, one.ctypes.data,
                                 one.ctypes.data, u.ctypes.data, 1, v.ctypes.data, 2,
                                 y.ctypes.data, None, 0) == 0
