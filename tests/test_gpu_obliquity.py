"""GPU parity of the obliquity-weighted superposition (HOLO_OBLIQUITY): every
pair also carries cos = |dz| / d, the `cosfact` of the MATLAB light-field
model (matlab/holofunctions.m:226-227).  Same bars as the unweighted path:
FP64 relative L2 <= 1e-12 against the oracle, FP32-split <= 1e-6.
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


def geom(seed, N, M, z, spread=4e-3, thick=0.0):
    g = np.random.default_rng(seed)
    u = np.zeros((N, 3))
    u[:, :2] = (g.random((N, 2)) - 0.5) * spread
    u[:, 2] = (g.random(N) - 0.5) * thick
    v = np.zeros((M, 3))
    v[:, :2] = (g.random((M, 2)) - 0.5) * spread
    v[:, 2] = z
    return g.random(N), g.random(N) * 2 * np.pi, u, v


@pytest.mark.parametrize("sign", [+1, -1])
@pytest.mark.parametrize("libm", [0, 1])
@pytest.mark.parametrize("z", [0.05, -0.3])
def test_fp64_obliquity_matches_oracle(pkg, ctx, oracle, sign, libm, z):
    """Targets above and below the sources (|dz| keeps the factor positive)."""
    pkg.set_option("f64_libm", libm)
    amp, phase, u, v = geom(40, 3000, 700, z)
    y = ctx.transform_full(amp, phase, u, v, sign, LAM, pkg.FP64, obliquity=True)
    ref = oracle.transform_full(sign, LAM, amp, phase, u, v, obliquity=True)
    assert rel_l2(y, ref) <= 1e-12
    y0 = ctx.transform_full(amp, phase, u, v, sign, LAM, pkg.FP64)
    assert rel_l2(y0, ref) > 1e-6  # the flag is not ignored


@pytest.mark.parametrize("R", [1, 2, 4])
def test_fp64_obliquity_lattice(pkg, ctx, oracle, workloads, R):
    """32^2 -> 64^2 lattice, every launch shape (targets per lane)."""
    pkg.set_option("r_f64", R)
    st = workloads.plain_small(0).stages[0]
    y = ctx.transform_full(st.amp, st.phase, st.u, st.v, +1, st.lam, pkg.FP64, obliquity=True)
    ref = oracle.transform_full(+1, st.lam, st.amp, st.phase, st.u, st.v, obliquity=True)
    assert rel_l2(y, ref) <= 1e-12


@pytest.mark.parametrize("thick, z", [(0.0, 0.3), (0.0, -0.2), (1e-6, 0.25), (1e-4, 0.3)])
def test_fp32_split_obliquity(pkg, ctx, oracle, thick, z):
    """Planar tiles (|dz| folded into the per-tile coefficients), thin and
    thick slabs (|dz| per pair; thick tiles go through the exact path)."""
    amp, phase, u, v = geom(41, 20000, 300, z, thick=thick)
    y = ctx.transform_full(amp, phase, u, v, +1, LAM, pkg.FP32_SPLIT, obliquity=True)
    ref = oracle.transform_full(+1, LAM, amp, phase, u, v, obliquity=True)
    assert rel_l2(y, ref) <= 1e-6


@pytest.mark.parametrize("z", [1e-3, 5e-3])
def test_fp32_split_obliquity_near_field(pkg, ctx, oracle, workloads, z):
    """A few mm from a 10 um lattice the cosine varies strongly across the
    field; the guard sends the tiles that need it through exact pairs."""
    u = workloads.grid(96, 96, 10e-6, 0.0)
    v = workloads.grid(40, 40, 25e-6, z)
    g = np.random.default_rng(42)
    amp, phase = g.random(u.shape[0]), g.random(u.shape[0]) * 2 * np.pi
    y = ctx.transform_full(amp, phase, u, v, -1, LAM, pkg.FP32_SPLIT, obliquity=True)
    ref = oracle.transform_full(-1, LAM, amp, phase, u, v, obliquity=True)
    assert rel_l2(y, ref) <= 1e-6


def test_plan_complex_weights_obliquity(pkg, ctx, oracle):
    amp, phase, u, v = geom(43, 2500, 900, 0.1, thick=1e-5)
    w = amp * np.exp(1j * phase)
    plan = ctx.plan(u)
    ref = oracle.transform_cartesian(+1, LAM, w, u, v, obliquity=True)
    assert rel_l2(plan.transform_complex(w, v, +1, LAM, pkg.FP64, obliquity=True), ref) <= 1e-12
    assert rel_l2(plan.transform_complex(w, v, +1, LAM, pkg.FP32_SPLIT, obliquity=True),
                  ref) <= 1e-6


@pytest.mark.parametrize("shared", [False, True])
def test_mc_obliquity_matches_oracle(pkg, ctx, oracle, shared):
    amp, phase, u, v = geom(44, 5000, 2500, 0.2)
    y = ctx.transform_mc(amp, phase, u, v, 257, 3, +1, LAM, pkg.FP64, shared=shared,
                         obliquity=True)
    sub = np.arange(0, 2500, 7)
    ref = oracle.transform_mc(+1, LAM, amp, phase, u, v, 257, 3, targets=sub, shared=shared,
                              obliquity=True)
    assert rel_l2(y[sub], ref) <= 1e-12
    y32 = ctx.transform_mc(amp, phase, u, v, 257, 3, +1, LAM, pkg.FP32_SPLIT, shared=shared,
                           obliquity=True)
    assert rel_l2(y32, y) <= 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_obliquity_matches_oracle(pkg, ctx, oracle, workloads, precision):
    amp, theta, u, _ = geom(45, 3000, 1, 0.0)
    v = workloads.grid(40, 30, 10e-6, 0.35)
    g0 = np.random.default_rng(46)
    g = g0.standard_normal(1200) + 1j * g0.standard_normal(1200)
    plan = ctx.plan(v, 16)
    ga, gt = plan.adjoint(g, amp, theta, u, +1, LAM, precision, obliquity=True)
    ra, rt = oracle.adjoint(+1, LAM, amp, theta, u, g, v, obliquity=True)
    bar = 1e-11 if precision == 0 else 1e-6
    assert rel_l2(ga, ra) <= bar and rel_l2(gt, rt) <= bar


def test_propagator_obliquity_gradient(pkg, ctx, oracle):
    """torch autograd through Propagator(obliquity=True) equals the oracle's
    adjoint of the same loss."""
    import torch
    from holo_amd.torch_ops import Propagator
    amp, theta, u, v = geom(47, 1500, 400, 0.08)
    dev = torch.device("cuda", 0)
    prop = Propagator(torch.from_numpy(u).to(dev), torch.from_numpy(v).to(dev), +1, LAM,
                      pkg.FP64, obliquity=True)
    a = torch.from_numpy(amp).to(dev).requires_grad_(True)
    t = torch.from_numpy(theta).to(dev).requires_grad_(True)
    z = prop(a, t)
    loss = (z[:, 0] ** 2 + z[:, 1] ** 2).sum()
    loss.backward()
    zr = oracle.transform_full(+1, LAM, amp, theta, u, v, obliquity=True)
    assert rel_l2(z.detach().cpu().numpy().view(np.complex128).ravel(), zr) <= 1e-12
    ra, rt = oracle.adjoint(+1, LAM, amp, theta, u, 2 * zr, v, obliquity=True)
    assert rel_l2(a.grad.cpu().numpy(), ra) <= 1e-11
    assert rel_l2(t.grad.cpu().numpy(), rt) <= 1e-11


def test_unknown_flag_rejected(pkg, ctx):
    L = pkg.lib()
    y = np.empty(2)
    one = np.ones(1)
    pos = np.array([0.0, 0.0, 0.0])
    tgt = np.array([0.0, 0.0, 0.1])
    st = L.holo_transform_full(ctx.handle, pkg.FP64, 1, LAM, one.ctypes.data, one.ctypes.data,
                               pos.ctypes.data, 1, tgt.ctypes.data, 1, y.ctypes.data, None,
                               0x40)
    assert st == 1 and b"flag" in L.holo_last_error()
    st = L.holo_mc_indices(ctx.handle, 0, 10, 0, 1, 1, y.ctypes.data, pkg.OBLIQUITY)
    assert st == 1
