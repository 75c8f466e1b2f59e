"""Per-target FP32-split Monte Carlo on product-grid sources.

When the sources form an x-slow product grid (the reference's plane
generators, cuda/init.h:89-158), the kernel takes x and y of a drawn source
from row/column tables and gathers only its 8-byte weight, instead of a
32-byte record.  The grid is recognised on the device; the pair inputs are
the same values, so the estimate must equal the generic gather's bit for bit
(option "mc_lattice" = 0), on and off the grid.
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


def weights(seed, N):
    g = np.random.default_rng(seed)
    return g.random(N), g.random(N) * 2 * np.pi


def targets(seed, M, z=0.3, half=3e-3):
    g = np.random.default_rng(seed)
    v = np.zeros((M, 3))
    v[:, :2] = (g.random((M, 2)) - 0.5) * 2 * half
    v[:, 2] = z
    return v


def both(pkg, ctx, amp, phase, u, v, K, sign=+1, obliquity=False):
    pkg.set_option("mc_lattice", 1)
    y1 = ctx.transform_mc(amp, phase, u, v, K, 7, sign, LAM, pkg.FP32_SPLIT, obliquity=obliquity)
    pkg.set_option("mc_lattice", 0)
    y0 = ctx.transform_mc(amp, phase, u, v, K, 7, sign, LAM, pkg.FP32_SPLIT, obliquity=obliquity)
    pkg.set_option("mc_lattice", 1)
    return y1, y0


def product_grid(xs, ys, z):
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], 1)


@pytest.mark.parametrize("shape", [(64, 48), (1, 300), (300, 1), (37, 29), (1024, 70)])
@pytest.mark.parametrize("sign", [+1, -1])
def test_mc_lattice_bitwise(pkg, ctx, workloads, shape, sign):
    u = workloads.grid(shape[0], shape[1], 10e-6, 0.0)
    amp, phase = weights(1, len(u))
    v = targets(2, 700)
    y1, y0 = both(pkg, ctx, amp, phase, u, v, 333, sign)
    assert np.array_equal(y1, y0)


@pytest.mark.parametrize("shape, pad", [((31, 32), 8), ((37, 29), 28), ((64, 48), 1)])
def test_mc_lattice_padded_plane(pkg, ctx, oracle, workloads, shape, pad):
    """A lattice followed by copies of source 0 (the reference's plane
    generator, cuda/init.h:153-157): the copies come from the packed records,
    the same bits as the generic gather, and they are drawn."""
    u = workloads.grid(shape[0], shape[1], 10e-6, 0.0)
    u = np.concatenate([u, np.repeat(u[:1], pad, 0)])
    amp, phase = weights(11, len(u))
    v = targets(12, 600)
    y1, y0 = both(pkg, ctx, amp, phase, u, v, 1024, -1)
    assert np.array_equal(y1, y0)
    amp2 = amp.copy()
    amp2[-pad:] = 0.0
    pkg.set_option("mc_lattice", 1)
    y2 = ctx.transform_mc(amp2, phase, u, v, 1024, 7, -1, LAM, pkg.FP32_SPLIT)
    assert not np.array_equal(y2, y1)


def test_mc_lattice_nonuniform_offcentre_obliquity(pkg, ctx):
    g = np.random.default_rng(3)
    u = product_grid(np.sort(g.random(40)) * 1e-3 + 2e-4, np.sort(g.random(24)) * 1e-3 - 1e-3, 0.01)
    amp, phase = weights(4, len(u))
    v = targets(5, 500, 0.25)
    y1, y0 = both(pkg, ctx, amp, phase, u, v, 257, obliquity=True)
    assert np.array_equal(y1, y0)


@pytest.mark.parametrize("case", ["random", "one_moved", "one_z", "row_too_long"])
def test_mc_off_grid_sources(pkg, ctx, workloads, case):
    """Point sets that are not product grids take the generic gather."""
    if case == "random":
        g = np.random.default_rng(6)
        u = np.zeros((3000, 3))
        u[:, :2] = (g.random((3000, 2)) - 0.5) * 2e-3
    elif case == "row_too_long":
        u = workloads.grid(2, 70000, 1e-6, 0.0)  # rows longer than the device scan
    else:
        u = workloads.grid(50, 40, 10e-6, 0.0)
        if case == "one_moved":
            u[1234, 1] += 1e-9
        else:
            u[999, 2] = 1e-7
    amp, phase = weights(7, len(u))
    v = targets(8, 300)
    y1, y0 = both(pkg, ctx, amp, phase, u, v, 129)
    assert np.array_equal(y1, y0)


def test_mc_lattice_matches_oracle(pkg, ctx, oracle, workloads):
    u = workloads.grid(96, 80, 10e-6, 0.0)
    amp, phase = weights(9, len(u))
    v = targets(10, 2000)
    y = ctx.transform_mc(amp, phase, u, v, 513, 11, +1, LAM, pkg.FP32_SPLIT)
    sub = np.arange(0, 2000, 9)
    ref = oracle.transform_mc(+1, LAM, amp, phase, u, v, 513, 11, targets=sub)
    assert rel_l2(y[sub], ref) <= 1e-6


# --------------------------------------------------------------------------
# per-target Monte Carlo in passes over source ranges (option mc_passes)


@pytest.mark.parametrize("passes", [2, 3, 7])
@pytest.mark.parametrize("case", ["lattice", "padded", "cloud"])
def test_mc_passes_match_oracle(pkg, ctx, oracle, workloads, passes, case):
    """Each pass takes the draws of one source range; together they are the
    estimator with the same indices: FP32-split within 1e-6 and FP64 within
    1e-12 of the oracle, and the table gather bit-identical to the record
    gather.  M = 700 leaves a partial wave; K = 1001 is odd and long enough to
    force partial batches of the per-lane queues."""
    if case == "lattice":
        u = workloads.grid(96, 80, 10e-6, 0.0)
    elif case == "padded":
        u = workloads.reference_plane(7000, 1.0e-3, 0.0, 16 / 9)
    else:
        g = np.random.default_rng(31)
        u = np.zeros((6000, 3))
        u[:, :2] = (g.random((6000, 2)) - 0.5) * 2e-3
    amp, phase = weights(passes, len(u))
    v = targets(32, 700)
    pkg.set_option("mc_passes", passes)
    K = 1001
    y1, y0 = both(pkg, ctx, amp, phase, u, v, K, -1)
    assert np.array_equal(y1, y0)
    sub = np.arange(0, 700, 7)
    ref = oracle.transform_mc(-1, LAM, amp, phase, u, v, K, 7, targets=sub)
    assert rel_l2(y1[sub], ref) <= 1e-6
    y64 = ctx.transform_mc(amp, phase, u, v, K, 7, -1, LAM, pkg.FP64)
    assert rel_l2(y64[sub], ref) <= 1e-12


@pytest.mark.parametrize("K", [1, 7, 8, 9, 64])
def test_mc_passes_ragged_k(pkg, ctx, oracle, workloads, K):
    """Few draws per target: the queues drain with partial batches only."""
    u = workloads.grid(64, 48, 10e-6, 0.0)
    amp, phase = weights(K, len(u))
    v = targets(33, 333)
    pkg.set_option("mc_passes", 4)
    y = ctx.transform_mc(amp, phase, u, v, K, 5, +1, LAM, pkg.FP32_SPLIT, obliquity=True)
    ref = oracle.transform_mc(+1, LAM, amp, phase, u, v, K, 5, obliquity=True)
    assert rel_l2(y, ref) <= 1e-6


def test_mc_passes_independent_of_count(pkg, ctx, workloads):
    """The estimate does not depend on the number of passes beyond FP32
    summation order (1 pass is the direct kernel)."""
    u = workloads.grid(128, 96, 10e-6, 0.0)
    amp, phase = weights(8, len(u))
    v = targets(34, 1500)
    ys = []
    for passes in (1, 2, 5, 16):
        pkg.set_option("mc_passes", passes)
        ys.append(ctx.transform_mc(amp, phase, u, v, 2048, 3, +1, LAM, pkg.FP32_SPLIT))
    for y in ys[1:]:
        assert rel_l2(y, ys[0]) <= 5e-7
