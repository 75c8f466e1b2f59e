"""Host-side lattice detection (holo_lattice_layout; no device).

A plan takes the lattice kernel when its sources are the planes of the
reference's generator (cuda/init.h:89-158): x = floor(sqrt(n * aspect)) rows
of y = n // x points, x slow, and the n - x*y points left over are copies of
point 0.  With an aspect ratio above 1 the leftovers can number y or more;
the detection must still find the x rows (a complete row has one x and row
0's y values, which the copies of point 0 never complete).
"""
import numpy as np
import pytest


def expected(n, aspect):
    x = int(np.sqrt(n * aspect))
    return x, n // x


@pytest.mark.parametrize("aspect", [16 / 9, 4 / 3, 1.0, 0.8, 9 / 16])
def test_generator_planes_detected_over_1000_sizes(pkg, workloads, aspect):
    """Every n in [20000, 21000): the lattice is (x, y), the n - x*y copies
    are the tail and 16-row tiles are built."""
    missed = []
    for n in range(20000, 21000):
        u = workloads.reference_plane(n, 1.2e-3, 0.0, aspect, (1e-4, -2e-4))
        x, y = expected(n, aspect)
        got = pkg.lattice_layout(u)
        if got != (x, y, n - x * y, True):
            missed.append((n, got))
    assert not missed, missed[:5]


def test_remainder_of_more_than_one_row():
    """n = 20999 at 16:9: 193 x 108 plus 155 copies of point 0 (more than one
    row of 108), the case the old nx = N / ny rule took for 194 rows."""
    n, aspect = 20999, 16 / 9
    x, y = expected(n, aspect)
    assert (x, y, n - x * y) == (193, 108, 155)


def test_detected_case(pkg, workloads):
    n = 20999
    u = workloads.reference_plane(n, 1.2e-3, 0.0, 16 / 9)
    assert pkg.lattice_layout(u) == (193, 108, 155, True)


def test_not_lattices(pkg, workloads):
    u = workloads.grid(64, 48, 10e-6, 0.0)
    assert pkg.lattice_layout(u) == (64, 48, 0, True)
    # a product grid with non-uniform x, sides multiples of 8: 8 x 8 tiles only
    g = np.random.default_rng(0)
    X, Y = np.meshgrid(np.sort(g.random(64)) * 1e-3, np.arange(48) * 1e-5, indexing="ij")
    un = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    assert pkg.lattice_layout(un) == (64, 48, 0, False)
    # non-uniform and 60 % 8 != 0: not a lattice
    X, Y = np.meshgrid(np.sort(g.random(60)) * 1e-3, np.arange(48) * 1e-5, indexing="ij")
    un = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    assert pkg.lattice_layout(un) == (0, 0, 0, False)
    assert pkg.lattice_layout(u[g.permutation(len(u))])[:2] == (0, 0)
    uz = u.copy()
    uz[5, 2] = 1e-9
    assert pkg.lattice_layout(uz)[:2] == (0, 0)
    # a tail of max(nx, ny) sources or more is not a padded lattice
    up = np.concatenate([workloads.grid(40, 36, 10e-6, 0.0), np.repeat(u[:1], 40, 0)])
    assert pkg.lattice_layout(up)[:2] == (0, 0)
    up = np.concatenate([workloads.grid(40, 36, 10e-6, 0.0), np.repeat(u[:1], 39, 0)])
    assert pkg.lattice_layout(up) == (40, 36, 39, True)
    # jittered planes (the reference's -r) have no product structure
    uj = u.copy()
    uj[:, :2] += (g.random((len(u), 2)) - 0.5) * 0.975 * 10e-6
    assert pkg.lattice_layout(uj)[:2] == (0, 0)
