"""Host-side tiling of source positions for the FP32-split kernels (no GPU)."""
import numpy as np
import pytest


def grid(n, pitch, z=0.0):
    i = (np.arange(n) - (n - 1) / 2) * pitch
    X, Y = np.meshgrid(i, i, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(n * n, z)], axis=1)


def check_partition(perm, N):
    used = perm[perm >= 0]
    assert np.array_equal(np.sort(used), np.arange(N))  # every source exactly once


@pytest.mark.parametrize("n, tile", [(1024, 64), (1024, 16), (512, 64), (32, 64), (64, 16)])
def test_lattice_gets_exact_blocks(pkg, n, tile):
    u = grid(n, 10e-6)
    perm, nt, off, planar = pkg.plan_layout(u, tile)
    check_partition(perm, n * n)
    assert nt == n * n // tile and np.all(perm >= 0)  # no padding
    b = int(round(np.sqrt(tile)))
    assert off == pytest.approx((b - 1) / 2 * 10e-6 * np.sqrt(2), rel=1e-6)
    assert planar
    # each tile is a b x b block of the lattice
    ij = np.stack(np.divmod(perm.reshape(nt, tile), n), axis=-1)
    span = ij.max(axis=1) - ij.min(axis=1)
    assert np.all(span == b - 1)


def test_random_plane_is_compact_and_mostly_full(pkg):
    g = np.random.default_rng(0)
    N = 50000
    u = np.zeros((N, 3))
    u[:, :2] = (g.random((N, 2)) - 0.5) * 4e-3
    perm, nt, off, planar = pkg.plan_layout(u, 64)
    check_partition(perm, N)
    assert planar
    assert N / (nt * 64) > 0.6          # padding stays bounded
    cell = np.sqrt(16e-6 * 64 / N)
    assert off < 2.5 * cell              # tiles stay within ~2 cells


def test_thin_slab_bins_in_two_dimensions(pkg):
    g = np.random.default_rng(1)
    N = 30000
    u = (g.random((N, 3)) - 0.5) * np.array([3e-3, 3e-3, 1e-6])
    perm, nt, off, planar = pkg.plan_layout(u, 64)
    check_partition(perm, N)
    assert not planar
    assert N / (nt * 64) > 0.6


def test_tiny_and_degenerate_inputs(pkg):
    perm, nt, off, planar = pkg.plan_layout(np.zeros((0, 3)), 16)
    assert nt == 0
    u = np.array([[1e-3, 2e-3, 0.3]])
    perm, nt, off, planar = pkg.plan_layout(u, 16)
    assert nt == 1 and perm[0] == 0 and np.all(perm[1:] == -1) and off == 0.0
    same = np.tile([[0.0, 0.0, 0.1]], (100, 1))  # all sources at one point
    perm, nt, off, planar = pkg.plan_layout(same, 16)
    check_partition(perm, 100)
    assert nt == 7 and off == 0.0


def test_invalid_tile(pkg):
    with pytest.raises(pkg.HoloError):
        pkg.plan_layout(grid(4, 1e-5), 12)
