"""The error report of every result (err_rel_l2 of holo.h).

An FP32-split call also evaluates the FP64 formulation at count = min(M, 256)
stride targets m_i = floor(i M / count) and reports
sqrt(sum |r[m_i] - r64_i|^2 / sum |r64_i|^2).  Each test recomputes that
number on the host from the oracle at the same targets and requires the two to
agree to 1e-3 of the value (the library's FP64 values are within 1e-12 of the
oracle's); FP64 calls report 0.  Host and device pointers both carry it.
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
    sub = tmap(M)
    ref = oracle.transform_full(-1, LAM, amp, phase, u, v, targets=sub)
    # transform_full, host pointers
    y = ctx.transform_full(amp, phase, u, v, -1, LAM, pkg.FP32_SPLIT, report_error=True)
    agree(ctx.last_rel_l2, rel_l2(y[sub], ref))
    # plan, device pointers (the report stays on the device)
    da, dp, du, dv = to_dev(amp, phase, u, v)
    plan = ctx.plan(du)
    yd = plan.transform(da, dp, dv, -1, LAM, pkg.FP32_SPLIT, report_error=True)
    e = ctx.last_rel_l2
    assert hasattr(e, "device") and e.device.type == "cuda"
    agree(float(e.item()), rel_l2(yd.cpu().numpy()[sub], ref))
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
    sub = tmap(len(v))
    ref = oracle.transform_cartesian(+1, LAM, w, u, v[sub], obliquity=True)
    agree(ctx.last_rel_l2, rel_l2(y[sub], ref))


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
    sub = tmap(len(v))
    ref = oracle.transform_mc(+1, LAM, amp, phase, u, v, K, 9, targets=sub, shared=shared)
    agree(ctx.last_rel_l2, rel_l2(y[sub], ref))
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
    sub = tmap(len(u))
    rb = oracle.transform_cartesian(-1, LAM, gz, v, u[sub])
    agree(ctx.last_rel_l2, rel_l2(b[sub], rb))


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
