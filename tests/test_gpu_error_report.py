"""The error report of every result (err_rel_l2 of holo.h).

An FP32-split call also evaluates the FP64 formulation at count = min(M, 256)
targets, half drawn by the energy of its result r and half spread evenly, and
reports the importance-weighted estimate of the whole result's relative L2
error (the host replica and its formula are in _report.py; every target once,
so the exact value, when M <= 256).  Each test replays the choice on the
host, evaluates the oracle at the chosen targets and requires the two reports
to agree to 1e-3 of the value (the library's FP64 values are within 1e-12 of
the oracle's); FP64 calls report 0.  Host and device pointers both carry it.
A focused field whose energy lies away from any stride subset checks that the
report follows the whole result's error.
"""
import sys

import numpy as np
import pytest

from _report import pick, report_of, tmap

pytestmark = pytest.mark.gpu
LAM = 632.8e-9


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.nodeid}", flush=True)
    sys.stdout.flush()
    yield


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


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
    pk = pick(y)
    ref = oracle.transform_full(-1, LAM, amp, phase, u, v, targets=pk[0])
    agree(ctx.last_rel_l2, report_of(y, pk, ref))
    # plan, device pointers (the report stays on the device)
    da, dp, du, dv = to_dev(amp, phase, u, v)
    plan = ctx.plan(du)
    yd = plan.transform(da, dp, dv, -1, LAM, pkg.FP32_SPLIT, report_error=True)
    e = ctx.last_rel_l2
    assert hasattr(e, "device") and e.device.type == "cuda"
    yd = yd.cpu().numpy()
    pkd = pick(yd)
    refd = oracle.transform_full(-1, LAM, amp, phase, u, v, targets=pkd[0])
    agree(float(e.item()), report_of(yd, pkd, refd))
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
    pk = pick(y)
    ref = oracle.transform_cartesian(+1, LAM, w, u, v[pk[0]], obliquity=True)
    agree(ctx.last_rel_l2, report_of(y, pk, ref))


@pytest.mark.parametrize("shared", [False, True])
@pytest.mark.parametrize("lattice", [False, True])
def test_monte_carlo(pkg, ctx, oracle, workloads, shared, lattice):
    """The FP64 estimator with the same indices at the sampled targets."""
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
    pk = pick(y)
    ref = oracle.transform_mc(+1, LAM, amp, phase, u, v, K, 9, targets=pk[0], shared=shared)
    agree(ctx.last_rel_l2, report_of(y, pk, ref))
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
    pk = pick(b)
    rb = oracle.transform_cartesian(-1, LAM, gz, v, u[pk[0]])
    agree(ctx.last_rel_l2, report_of(b, pk, rb))


def test_focused_field_report_follows_whole_result(pkg, ctx, oracle, workloads):
    """Sources phased to focus near target 1, every 8th target (a stride
    subset of 2048) 2 mm to the side in the dark: the report stays within a
    factor 2 of the whole result's relative L2 error against the oracle and
    draws its energy samples from the bright targets, while the stride subset's
    own relative error would be far higher."""
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
    pk = pick(y)
    assert pk[1] and np.mean(pk[0][:pk[2]] % 8 == 0) < 0.05
    agree(rep, report_of(y, pk, full[pk[0]]))
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
