"""holo_transform_full_noalloc: the FP64 transform on device pointers with its
scratch in a static device array (no allocation).  Called through the C ABI
(ctypes) on torch's current stream and compared with holo_transform_full on
the same device inputs, and with the oracle, to the FP64 bar (relative L2
<= 1e-12).  Two calls in a row, the second with a smaller input, so that the
second runs on the array the first left behind.
"""
import ctypes
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


def noalloc(pkg, amp, phase, u, v, sign, lam, flags=0):
    """y = holo_transform_full_noalloc(...) on torch's current stream."""
    import torch
    N, M = amp.numel(), v.numel() // 3
    y = torch.empty((M, 2), dtype=torch.float64, device=v.device)
    stream = torch.cuda.current_stream(v.device).cuda_stream
    st = pkg.lib().holo_transform_full_noalloc(sign, lam, amp.data_ptr() if N else None,
                                               phase.data_ptr() if N else None,
                                               u.data_ptr() if N else None, N, v.data_ptr(),
                                               M, y.data_ptr(), stream, flags)
    pkg.check(st)
    return torch.view_as_complex(y).cpu().numpy()


def geom(seed, N, M, planar):
    g = np.random.default_rng(seed)
    u = np.zeros((N, 3))
    u[:, :2] = (g.random((N, 2)) - 0.5) * 4e-3
    if not planar:
        u[:, 2] = g.random(N) * 1e-4
    v = np.zeros((M, 3))
    v[:, :2] = (g.random((M, 2)) - 0.5) * 4e-3
    v[:, 2] = 0.3
    return g.random(N), g.random(N) * 2 * np.pi, u, v


def test_two_calls_second_smaller(pkg, ctx, oracle):
    import torch
    dev = torch.device("cuda", 0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    for seed, N, M, planar, sign in ((1, 4096, 20000, False, +1), (2, 1000, 3000, True, -1)):
        amp, phase, u, v = geom(seed, N, M, planar)
        y = noalloc(pkg, T(amp), T(phase), T(u), T(v), sign, LAM)
        ref_full = ctx.transform_full(T(amp), T(phase), T(u), T(v), sign, LAM,
                                      pkg.FP64).cpu().numpy()
        assert rel_l2(y, ref_full) <= 1e-12, (N, M)
        sub = np.arange(0, M, 17)
        ref = oracle.transform_full(sign, LAM, amp, phase, u, v, targets=sub)
        assert rel_l2(y[sub], ref) <= 1e-12, (N, M)


def test_lattice_rows_and_obliquity(pkg, ctx, oracle, workloads):
    """A padded lattice (rows of one x in one plane: the layout words in the
    static array select the rows path) with obliquity, then a non-planar
    cloud on the same array: both as holo_transform_full and the oracle."""
    import torch
    dev = torch.device("cuda", 0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    u = workloads.reference_plane(20999, 1.2e-3, 0.0, 16 / 9)
    g = np.random.default_rng(9)
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    v = workloads.grid(64, 48, 7.56e-6, 0.25)
    y = noalloc(pkg, T(amp), T(phase), T(u), T(v), +1, LAM, pkg.OBLIQUITY)
    ref_full = ctx.transform_full(T(amp), T(phase), T(u), T(v), +1, LAM, pkg.FP64,
                                  obliquity=True).cpu().numpy()
    assert rel_l2(y, ref_full) <= 1e-12
    sub = np.arange(0, len(v), 29)
    ref = oracle.transform_full(+1, LAM, amp, phase, u, v, targets=sub, obliquity=True)
    assert rel_l2(y[sub], ref) <= 1e-12
    amp2, phase2, u2, v2 = geom(4, 777, 500, False)
    y2 = noalloc(pkg, T(amp2), T(phase2), T(u2), T(v2), -1, LAM)
    ref2 = oracle.transform_full(-1, LAM, amp2, phase2, u2, v2)
    assert rel_l2(y2, ref2) <= 1e-12


def test_large_target_count_one_split(pkg, ctx, workloads):
    """M = 2^21 targets: more than the array holds for two splits, so the call
    takes one split and writes y directly; same field as holo_transform_full
    (which may split) to FP64 rounding."""
    import torch
    dev = torch.device("cuda", 0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    u = workloads.grid(32, 32, 10e-6, 0.0)
    g = np.random.default_rng(10)
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    v = workloads.grid(2048, 1024, 7.56e-6, 0.35)
    y = noalloc(pkg, T(amp), T(phase), T(u), T(v), +1, LAM)
    ref = ctx.transform_full(T(amp), T(phase), T(u), T(v), +1, LAM, pkg.FP64).cpu().numpy()
    assert rel_l2(y, ref) <= 1e-12


def test_rejects_bad_arguments(pkg):
    import torch
    dev = torch.device("cuda", 0)
    a = torch.zeros(4, dtype=torch.float64, device=dev)
    u = torch.zeros(12, dtype=torch.float64, device=dev)
    y = torch.empty(8, dtype=torch.float64, device=dev)
    L = pkg.lib()
    s = torch.cuda.current_stream(dev).cuda_stream
    assert L.holo_transform_full_noalloc(0, LAM, a.data_ptr(), a.data_ptr(), u.data_ptr(), 4,
                                         u.data_ptr(), 4, y.data_ptr(), s, 0) == 1
    assert L.holo_transform_full_noalloc(1, -1.0, a.data_ptr(), a.data_ptr(), u.data_ptr(), 4,
                                         u.data_ptr(), 4, y.data_ptr(), s, 0) == 1
    assert L.holo_transform_full_noalloc(1, LAM, a.data_ptr(), a.data_ptr(), u.data_ptr(), 4,
                                         u.data_ptr(), 4, y.data_ptr(), s, pkg.MC_SHARED) == 1
    assert L.holo_transform_full_noalloc(1, LAM, a.data_ptr(), a.data_ptr(), u.data_ptr(), 4,
                                         None, 4, y.data_ptr(), s, 0) == 1
    assert b"flag" in L.holo_last_error() or b"null" in L.holo_last_error()
