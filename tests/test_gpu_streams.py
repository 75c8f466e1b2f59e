"""Stream semantics of HOLO_DEVICE_POINTERS calls (include/holo.h).

* every call runs on torch's current stream: a side stream gives the same
  bits as the default stream (FP64, FP32-split lattice, adjoint, Monte Carlo);
* a context that switches streams between calls orders its shared scratch
  (the FP64 split partials) on the device: two back-to-back calls on two
  streams give the bits each gives alone;
* a plan that serves two contexts on two streams orders its weight buffers
  the same way.

All streams are created at the default priority.
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


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _same(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _lattice(workloads, n_src, n_tgt, seed):
    g = np.random.default_rng(seed)
    u = workloads.grid(n_src, n_src, 10e-6, 0.0)
    v = workloads.grid(n_tgt, n_tgt, 7.56e-6, 0.35)
    return g.random(n_src * n_src), g.random(n_src * n_src) * 2 * np.pi, u, v


def test_side_stream_same_bits(pkg, ctx, workloads):
    import torch
    amp, phase, u, v = (_dev(a) for a in _lattice(workloads, 128, 64, 1))
    g = _dev(np.random.default_rng(2).standard_normal((64 * 64, 2)))
    plan_v = ctx.plan(v)

    def run():
        plan_u = ctx.plan(u)
        return [ctx.transform_full(amp, phase, u, v, +1, LAM, pkg.FP64),
                plan_u.transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT),
                *plan_v.adjoint(g, amp, phase, u, +1, LAM, pkg.FP32_SPLIT),
                ctx.transform_mc(amp, phase, u, v, 512, 7, +1, LAM, pkg.FP32_SPLIT)]

    ref = run()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        out = run()
    torch.cuda.synchronize()
    assert plan_v.lattice() == (64, 64)
    for a, b in zip(out, ref):
        assert _same(a, b)


def test_stream_switch_orders_context_scratch(pkg, ctx):
    """Two FP64 calls that split over sources share the context's partials."""
    import torch
    g = np.random.default_rng(3)
    N, M = 1 << 21, 2048
    u = np.zeros((N, 3))
    u[:, :2] = (g.random((N, 2)) - 0.5) * 5e-3
    v = np.zeros((M, 3))
    v[:, :2] = (g.random((M, 2)) - 0.5) * 5e-3
    v[:, 2] = 0.35
    args = [(_dev(g.random(N)), _dev(g.random(N) * 2 * np.pi)) for _ in range(2)]
    du, dv = _dev(u), _dev(v)
    ref = []
    for a, p in args:
        ref.append(ctx.transform_full(a, p, du, dv, +1, LAM, pkg.FP64))
        torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    s1.wait_stream(torch.cuda.current_stream())
    s2.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s1):
        y1 = ctx.transform_full(*args[0], du, dv, +1, LAM, pkg.FP64)
    with torch.cuda.stream(s2):
        y2 = ctx.transform_full(*args[1], du, dv, +1, LAM, pkg.FP64)
    torch.cuda.synchronize()
    assert _same(y1, ref[0]) and _same(y2, ref[1])


def test_plan_shared_by_two_contexts(pkg, workloads):
    """The lattice plan's weight buffers, rewritten by each call, are ordered
    across two contexts on two streams."""
    import torch
    lib = pkg.lib()
    amp0, ph0, u, v = _lattice(workloads, 512, 256, 4)
    g = np.random.default_rng(5)
    weights = [(_dev(amp0), _dev(ph0)), (_dev(g.random(amp0.size)),
                                         _dev(g.random(amp0.size) * 2 * np.pi))]
    dv = _dev(v)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    c1 = pkg.Context(0, s1.cuda_stream)
    c2 = pkg.Context(0, s2.cuda_stream)
    try:
        plan = c1.plan(u)
        assert plan.lattice() == (512, 512)
        torch.cuda.synchronize()

        def call(c, w, y):
            pkg.check(lib.holo_plan_transform(c.handle, plan._h, pkg.FP32_SPLIT, +1, LAM,
                                              w[0].data_ptr(), w[1].data_ptr(),
                                              dv.data_ptr(), ctypes.c_size_t(v.shape[0]),
                                              y.data_ptr(), None, pkg.DEVICE_POINTERS))

        ys = [torch.empty((v.shape[0], 2), dtype=torch.float64, device="cuda")
              for _ in range(4)]
        call(c1, weights[0], ys[0])
        c1.synchronize()
        call(c1, weights[1], ys[1])
        c1.synchronize()
        torch.cuda.synchronize()
        call(c1, weights[0], ys[2])  # on s1
        call(c2, weights[1], ys[3])  # on s2, rewrites the plan's weights
        torch.cuda.synchronize()
        assert _same(ys[2], ys[0]) and _same(ys[3], ys[1])
        assert not _same(ys[0], ys[1])
        plan.close()
    finally:
        c1.close()
        c2.close()
