"""The matrix-core lattice kernels (HOLO_FP32_SPLIT on product-grid sources).

Every test runs four times: with options "grid16", "grid32" and "grid64" on
(uniform lattices use 16-row tiles, padded to multiples of 16, of 64 columns
where the tile columns form quads, else of 32 where they pair up, else of 16),
with "grid64" off (16 x 32 tiles where the columns pair up), with "grid32" off
(16 x 16 tiles), and with "grid16" off (8 x 8 tiles on product grids with
sides multiples of 8, the tiled kernel on other planes).
Bars: relative L2 <= 1e-6 against the FP64 oracle, as for every FP32-split
path; the lattice and the tiled kernel agree to the same bar; results do not
depend on the launch split or on the weights' scale.
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


@pytest.fixture(autouse=True, params=[64, 32, 16, 8],
                ids=["tiles64", "tiles32", "tiles16", "tiles8"])
def tiles(request, pkg):
    pkg.set_option("grid16", 0 if request.param == 8 else 1)
    pkg.set_option("grid32", 1 if request.param >= 32 else 0)
    pkg.set_option("grid64", 1 if request.param == 64 else 0)
    yield request.param
    pkg.set_option("grid16", 1)
    pkg.set_option("grid32", 1)
    pkg.set_option("grid64", 1)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def weights(seed, N):
    g = np.random.default_rng(seed)
    return g.random(N), g.random(N) * 2 * np.pi


def targets(seed, M, z, half=3e-3):
    g = np.random.default_rng(seed)
    v = np.zeros((M, 3))
    v[:, :2] = (g.random((M, 2)) - 0.5) * 2 * half
    v[:, 2] = z
    return v


def test_lattice_detection(pkg, ctx, workloads):
    assert ctx.plan(workloads.grid(64, 48, 10e-6, 0.0)).lattice() == (64, 48)
    # 60 is not a multiple of 8: a uniform lattice, padded for the 16-row tiles
    assert ctx.plan(workloads.grid(60, 48, 10e-6, 0.0)).lattice() == (60, 48)
    g = np.random.default_rng(0)
    xs = np.sort(g.random(60)) * 1e-3
    X, Y = np.meshgrid(xs, np.arange(48) * 1e-5, indexing="ij")
    un = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    assert ctx.plan(un).lattice() == (0, 0)  # not uniform, 60 % 8
    u = workloads.grid(64, 48, 10e-6, 0.0)
    assert ctx.plan(u[np.random.default_rng(0).permutation(len(u))]).lattice() == (0, 0)
    ux = u.reshape(64, 48, 3).transpose(1, 0, 2).reshape(-1, 3)  # x fast
    assert ctx.plan(ux).lattice() == (0, 0)
    uz = u.copy()
    uz[5, 2] = 1e-9
    assert ctx.plan(uz).lattice() == (0, 0)
    assert ctx.plan(u, 16).lattice() == (0, 0)  # the lattice kernel uses 8 x 8 tiles


@pytest.mark.parametrize("sign", [+1, -1])
@pytest.mark.parametrize("obliquity", [False, True])
def test_lattice_matches_oracle(pkg, ctx, oracle, workloads, sign, obliquity):
    u = workloads.grid(128, 96, 10e-6, 0.0)
    amp, phase = weights(1, len(u))
    v = targets(2, 777, 0.3)
    plan = ctx.plan(u)
    assert plan.lattice() == (128, 96)
    y = plan.transform(amp, phase, v, sign, LAM, pkg.FP32_SPLIT, obliquity=obliquity)
    ref = oracle.transform_full(sign, LAM, amp, phase, u, v, obliquity=obliquity)
    assert rel_l2(y, ref) <= 1e-6


def test_lattice_nonuniform_product_grid(pkg, ctx, oracle):
    """Any x_i by y_j product (not only uniform pitch), off-centre, z != 0."""
    g = np.random.default_rng(3)
    xs = np.sort(g.random(64)) * 1.2e-3 + 2e-4
    ys = np.sort(g.random(40)) * 0.8e-3 - 1e-3
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    u = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 0.01)], 1)
    amp, phase = weights(4, len(u))
    v = targets(5, 500, 0.26)
    plan = ctx.plan(u)
    assert plan.lattice() == (64, 40)
    y = plan.transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(y, oracle.transform_full(+1, LAM, amp, phase, u, v)) <= 1e-6


def test_lattice_and_tiled_kernels_agree(pkg, ctx, workloads):
    u = workloads.grid(256, 256, 10e-6, 0.0)
    amp, phase = weights(6, len(u))
    v = workloads.grid(48, 48, 7.56e-6, 0.35)
    plan = ctx.plan(u)
    y64 = plan.transform(amp, phase, v, +1, LAM, pkg.FP64)
    yg = plan.transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT)
    pkg.set_option("grid", 0)
    yt = plan.transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT)
    pkg.set_option("grid", 1)
    assert rel_l2(yg, y64) <= 1e-6 and rel_l2(yt, y64) <= 1e-6
    assert rel_l2(yg, yt) <= 1e-6


@pytest.mark.parametrize("splits", [1, 3, 7])
def test_lattice_split_independent(pkg, ctx, workloads, splits):
    u = workloads.grid(96, 96, 10e-6, 0.0)
    amp, phase = weights(7, len(u))
    v = targets(8, 300, 0.35)
    plan = ctx.plan(u)
    y0 = plan.transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    pkg.set_option("splits", splits)
    y1 = plan.transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    # the FP32 partial sums (flushed to FP64 every 8 tiles) regroup with the
    # split, so results move at the FP32 rounding level only
    assert rel_l2(y1, y0) <= 1e-6


@pytest.mark.parametrize("M", [1, 63, 64, 65, 257])
def test_lattice_ragged_target_counts(pkg, ctx, oracle, workloads, M):
    u = workloads.grid(64, 64, 10e-6, 0.0)
    amp, phase = weights(9, len(u))
    v = targets(10, M, 0.35)
    y = ctx.plan(u).transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT)
    assert y.shape == (M,)
    assert rel_l2(y, oracle.transform_full(+1, LAM, amp, phase, u, v)) <= 1e-6


@pytest.mark.parametrize("scale", [1e-30, 1e-8, 1e25])
def test_lattice_weight_scale(pkg, ctx, workloads, scale):
    """The f16 correction operand is scaled per tile: any weight scale works."""
    u = workloads.grid(64, 64, 10e-6, 0.0)
    amp, phase = weights(11, len(u))
    v = targets(12, 200, 0.35)
    plan = ctx.plan(u)
    y1 = plan.transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT)
    ys = plan.transform(amp * scale, phase, v, +1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(ys / scale, y1) <= 1e-6


@pytest.mark.parametrize("z", [1e-3, 5e-3, 2e-2])
def test_lattice_near_field_guard(pkg, ctx, oracle, workloads, z):
    """Near the source plane the separable expansion fails its bounds; those
    (target, tile) pairs run exact FP64-distance pairs inside the kernel."""
    u = workloads.grid(96, 96, 10e-6, 0.0)
    amp, phase = weights(13, len(u))
    v = workloads.grid(40, 40, 25e-6, z)
    y = ctx.plan(u).transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(y, oracle.transform_full(-1, LAM, amp, phase, u, v)) <= 1e-6


def test_lattice_complex_weights_and_adjoint(pkg, ctx, oracle, workloads):
    """Cartesian weights (the adjoint's sources) on a lattice plan, and the
    adjoint of a lattice forward plane."""
    v = workloads.grid(40, 32, 10e-6, 0.35)
    g0 = np.random.default_rng(14)
    gz = g0.standard_normal(len(v)) + 1j * g0.standard_normal(len(v))
    amp, theta = weights(15, 2500)
    u = targets(16, 2500, 0.0, 2e-3)
    plan = ctx.plan(v)
    assert plan.lattice() == (40, 32)
    ga, gt = plan.adjoint(gz, amp, theta, u, +1, LAM, pkg.FP32_SPLIT)
    ra, rt = oracle.adjoint(+1, LAM, amp, theta, u, gz, v)
    assert rel_l2(ga, ra) <= 1e-6 and rel_l2(gt, rt) <= 1e-6
    yc = plan.transform_complex(gz, u, -1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(yc, oracle.transform_cartesian(-1, LAM, gz, v, u)) <= 1e-6


def test_lattice_64_32_16_and_8_tiles_are_different_paths(pkg, ctx, oracle, workloads, tiles):
    """A uniform 128 x 128 lattice runs the 16 x 64 kernel with "grid16",
    "grid32" and "grid64" on, the 16 x 32 kernel with "grid64" off, the 16 x 16
    kernel with "grid32" off and the 8 x 8 kernel with "grid16" off: all four
    meet the bar and round differently."""
    u = workloads.grid(128, 128, 10e-6, 0.0)
    amp, phase = weights(17, len(u))
    v = targets(18, 500, 0.35, 4e-3)
    plan = ctx.plan(u)
    ref = oracle.transform_full(+1, LAM, amp, phase, u, v)
    ys = {}
    for name, g16, g32, g64 in (("64", 1, 1, 1), ("32", 1, 1, 0), ("16", 1, 0, 0),
                                ("8", 0, 1, 1)):
        pkg.set_option("grid16", g16)
        pkg.set_option("grid32", g32)
        pkg.set_option("grid64", g64)
        ys[name] = plan.transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT)
        assert rel_l2(ys[name], ref) <= 1e-6, name
    names = list(ys)
    for i, a in enumerate(names):
        for b in names[i + 1:]:
            assert 0 < rel_l2(ys[a], ys[b]) <= 1e-6, (a, b)


@pytest.mark.parametrize("sign, obliquity", [(+1, False), (-1, True)])
def test_lattice_16x32_waves_fall_back_to_16x16(pkg, ctx, oracle, workloads, sign, obliquity,
                                                tiles):
    """Targets between the two tile sizes' far-field bounds (z = 0.1 m, up to
    0.5 mm off axis): every wave fails the 16 x 32 bounds somewhere and takes
    the 16 x 16 halves of the same tiles, none of which needs the exact path.
    With amplitudes 0.75 every tile's largest weight component lies in
    [0.53, 0.75], so tiles and pairs alike have scale 1, and the result is the
    16 x 16 kernel's, bit for bit."""
    u = workloads.grid(64, 64, 10e-6, 0.0)
    amp = np.full(len(u), 0.75)
    phase = np.random.default_rng(25).random(len(u)) * 2 * np.pi
    v = targets(26, 512, 0.1, 5e-4)
    plan = ctx.plan(u)
    ys = {}
    for g32 in (1, 0):
        pkg.set_option("grid16", 1)
        pkg.set_option("grid32", g32)
        ys[g32] = plan.transform(amp, phase, v, sign, LAM, pkg.FP32_SPLIT, obliquity=obliquity)
    assert np.array_equal(ys[1], ys[0])
    ref = oracle.transform_full(sign, LAM, amp, phase, u, v, obliquity=obliquity)
    assert rel_l2(ys[1], ref) <= 1e-6


@pytest.mark.parametrize("nx, ny", [(16, 32), (16, 16), (32, 32), (8, 8), (48, 96), (16, 64),
                                    (48, 128)])
@pytest.mark.parametrize("M", [1, 65])
def test_lattice_smallest_tilings(pkg, ctx, oracle, workloads, nx, ny, M, tiles):
    """One 16 x 32 pair, one 16 x 16 tile, two pairs, one 8 x 8 tile, an odd
    number of tile rows, one 16 x 64 quad and two quads per row of tiles,
    against the oracle for one and for 65 targets."""
    u = workloads.grid(nx, ny, 10e-6, 0.0)
    amp, phase = weights(27, len(u))
    v = targets(28, M, 0.3, 2e-3)
    plan = ctx.plan(u)
    assert plan.lattice() == (nx, ny)
    y = plan.transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(y, oracle.transform_full(-1, LAM, amp, phase, u, v)) <= 1e-6


@pytest.mark.parametrize("sign, obliquity", [(+1, False), (-1, True)])
def test_lattice_16x64_waves_fall_back_to_16x32(pkg, ctx, oracle, workloads, sign, obliquity,
                                                tiles):
    """Targets within 0.1 mm of the axis of a 64 x 256 lattice at 0.15 m pass
    the 16 x 32 far-field bounds (at most 0.83 of each) but fail the 16 x 64
    ones (the dropped share of the cross term, at least 1.45 of its bound):
    every wave takes the 16 x 32 tiles of the same sources.  With amplitudes
    0.75 quads, pairs and tiles alike have weight scale 1, and the result is
    the 16 x 32 kernel's, bit for bit."""
    u = workloads.grid(64, 256, 10e-6, 0.0)
    amp = np.full(len(u), 0.75)
    phase = np.random.default_rng(29).random(len(u)) * 2 * np.pi
    v = targets(30, 512, 0.15, 1e-4)
    plan = ctx.plan(u)
    ys = {}
    for g64 in (1, 0):
        pkg.set_option("grid16", 1)
        pkg.set_option("grid32", 1)
        pkg.set_option("grid64", g64)
        ys[g64] = plan.transform(amp, phase, v, sign, LAM, pkg.FP32_SPLIT, obliquity=obliquity)
    assert np.array_equal(ys[1], ys[0])
    ref = oracle.transform_full(sign, LAM, amp, phase, u, v, obliquity=obliquity)
    assert rel_l2(ys[1], ref) <= 1e-6


@pytest.mark.parametrize("quarter, scale", [(0, 1e-30), (3, 1e-30), (2, 0.0)])
def test_lattice_tile_quad_weight_scales(pkg, ctx, oracle, workloads, quarter, scale, tiles):
    """The four 16 x 16 quarters of a 16 x 64 tile share one weight scale: one
    quarter's weights far below (or zero beside) the others' still meet the
    bar."""
    u = workloads.grid(32, 128, 10e-6, 0.0)
    amp, phase = weights(31, len(u))
    a2 = amp.reshape(32, 128).copy()
    for c0 in (0, 64):  # that quarter of both 16 x 64 tiles of each row of tiles
        a2[:, c0 + 16 * quarter:c0 + 16 * quarter + 16] *= scale
    amp = a2.reshape(-1)
    v = targets(32, 300, 0.3, 2e-3)
    y = ctx.plan(u).transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(y, oracle.transform_full(-1, LAM, amp, phase, u, v)) <= 1e-6


@pytest.mark.parametrize("half, scale", [(0, 1e-30), (1, 1e-30), (1, 0.0)])
def test_lattice_tile_pair_weight_scales(pkg, ctx, oracle, workloads, half, scale, tiles):
    """The two 16 x 16 halves of a 16 x 32 tile share one weight scale: one
    half's weights far below (or zero beside) the other's still meet the bar."""
    u = workloads.grid(32, 64, 10e-6, 0.0)
    amp, phase = weights(23, len(u))
    a2 = amp.reshape(32, 64).copy()
    for c0 in (0, 32):  # the first or second half of both 16 x 32 tiles of each row
        a2[:, c0 + 16 * half:c0 + 16 * half + 16] *= scale
    amp = a2.reshape(-1)
    v = targets(24, 300, 0.3, 2e-3)
    y = ctx.plan(u).transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(y, oracle.transform_full(-1, LAM, amp, phase, u, v)) <= 1e-6


@pytest.mark.parametrize("half, z", [(2e-2, 0.1), (5e-2, 0.2), (8e-3, 0.05)])
def test_lattice_steep_angles(pkg, ctx, oracle, workloads, half, z, tiles):
    """Targets far off axis: the phase advances by up to several cycles per
    source pitch, so it spans tens of cycles across a tile (the 16 x 16
    kernel reduces that linear part exactly)."""
    u = workloads.grid(64, 48, 10e-6, 0.0)
    amp, phase = weights(19, len(u))
    v = targets(20, 400, z, half)
    y = ctx.plan(u).transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(y, oracle.transform_full(-1, LAM, amp, phase, u, v)) <= 1e-6


@pytest.mark.parametrize("flip", ["x", "y", "xy"])
def test_lattice_descending_pitch(pkg, ctx, oracle, workloads, flip, tiles):
    """Uniform lattices traversed with decreasing x and/or y (negative pitch),
    off-centre and at z != 0."""
    u = workloads.grid(64, 32, 8e-6, 0.02).reshape(64, 32, 3)
    if "x" in flip:
        u = u[::-1]
    if "y" in flip:
        u = u[:, ::-1]
    u = np.ascontiguousarray(u.reshape(-1, 3)) + np.array([3e-4, -2e-4, 0.0])
    amp, phase = weights(21, len(u))
    v = targets(22, 300, 0.3)
    plan = ctx.plan(u)
    assert plan.lattice() == (64, 32)
    y = plan.transform(amp, phase, v, +1, LAM, pkg.FP32_SPLIT, obliquity=True)
    ref = oracle.transform_full(+1, LAM, amp, phase, u, v, obliquity=True)
    assert rel_l2(y, ref) <= 1e-6


def _random_lattice_case(seed):
    """A seeded random lattice geometry: sides, pitches (not square), lateral
    offsets, source z, target distance (3 cm to 1 m, log-uniform), wavelength,
    sign and obliquity."""
    g = np.random.default_rng(1000 + seed)
    nx = 16 * int(g.integers(1, 11))
    ny = (32 if g.random() < 0.5 else 16) * int(g.integers(1, 6))
    px, py = g.uniform(2e-6, 20e-6, 2)
    ox, oy = g.uniform(-2e-3, 2e-3, 2)
    z0 = g.uniform(-1e-2, 1e-2)
    x = (np.arange(nx) - (nx - 1) / 2.0) * px + ox
    y = (np.arange(ny) - (ny - 1) / 2.0) * py + oy
    X, Y = np.meshgrid(x, y, indexing="ij")
    u = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z0)], 1)
    z = z0 + float(np.exp(g.uniform(np.log(0.03), np.log(1.0))))
    spread = g.uniform(1e-3, 0.05) * (z - z0)
    v = np.zeros((int(g.integers(200, 1500)), 3))
    v[:, :2] = (g.random((len(v), 2)) - 0.5) * 2 * spread + g.uniform(-spread, spread, 2)
    v[:, 2] = z
    lam = g.uniform(400e-9, 700e-9)
    sign = 1 if g.random() < 0.5 else -1
    obl = bool(g.random() < 0.3)
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    return u, (nx, ny), amp, phase, v, lam, sign, obl


@pytest.mark.parametrize("seed", range(16))
def test_lattice_random_geometries(pkg, ctx, seed, tiles):
    """Seeded random lattices and target sets across distances, pitches,
    wavelengths and angles: FP32-split within 1e-6 of FP64 everywhere (the
    accuracy guard sends what the far-field expansion cannot hold to the exact
    path)."""
    u, shape, amp, phase, v, lam, sign, obl = _random_lattice_case(seed)
    plan = ctx.plan(u)
    assert plan.lattice() == shape
    y64 = plan.transform(amp, phase, v, sign, lam, pkg.FP64, obliquity=obl)
    y32 = plan.transform(amp, phase, v, sign, lam, pkg.FP32_SPLIT, obliquity=obl)
    assert rel_l2(y32, y64) <= 1e-6
    if tiles == 32:
        # the tiled kernel on the same lattice, and on the lattice with every
        # source moved by up to half a pitch (the reference's -r jitter)
        pkg.set_option("grid", 0)
        yt = plan.transform(amp, phase, v, sign, lam, pkg.FP32_SPLIT, obliquity=obl)
        pkg.set_option("grid", 1)
        assert rel_l2(yt, y64) <= 1e-6
        g = np.random.default_rng(seed)
        du = np.diff(np.unique(u[:, 0]))[0], np.diff(np.unique(u[:, 1]))[0]
        uj = u.copy()
        uj[:, 0] += (g.random(len(u)) - 0.5) * 0.975 * du[0]
        uj[:, 1] += (g.random(len(u)) - 0.5) * 0.975 * du[1]
        pj = ctx.plan(uj)
        assert pj.lattice() == (0, 0)
        yj64 = pj.transform(amp, phase, v, sign, lam, pkg.FP64, obliquity=obl)
        yj32 = pj.transform(amp, phase, v, sign, lam, pkg.FP32_SPLIT, obliquity=obl)
        assert rel_l2(yj32, yj64) <= 1e-6


@pytest.mark.parametrize("n, aspect", [(1000, 1.0), (61 * 45, 61 / 45), (1024 * 1080 // 64, 1.0),
                                       (40 * 50 + 17, 0.8)])
@pytest.mark.parametrize("sign", [+1, -1])
def test_lattice_padded_planes(pkg, ctx, oracle, workloads, n, aspect, sign, tiles):
    """Planes of the reference's generator whose sides are not multiples of 16,
    ending with the copies of point 0 that fill n: 16-row tiles over the
    lattice padded with weight-0 slots, the copies summed exactly (with
    "grid16" off: the tiled kernel)."""
    u = workloads.reference_plane(n, 1.2e-3, 0.0, aspect, (1e-4, -2e-4))
    x = int(np.sqrt(n * aspect))
    y = n // x
    g = np.random.default_rng(n)
    amp, phase = g.random(n), g.random(n) * 2 * np.pi
    v = targets(n % 97, 500, 0.3)
    plan = ctx.plan(u)
    assert plan.lattice() == (x, y)
    y32 = plan.transform(amp, phase, v, sign, LAM, pkg.FP32_SPLIT, obliquity=(sign < 0))
    ref = oracle.transform_full(sign, LAM, amp, phase, u, v, obliquity=(sign < 0))
    assert rel_l2(y32, ref) <= 1e-6
    # the copies of point 0 are sources too
    amp2 = amp.copy()
    amp2[x * y:] = 0.0
    y2 = plan.transform(amp2, phase, v, sign, LAM, pkg.FP32_SPLIT, obliquity=(sign < 0))
    if n > x * y:
        assert rel_l2(y2, y32) > 1e-4


# near-field paths of the 16-row kernels: (grid_near, grid_series)
NEAR_PATHS = ((1, 1), (0, 1), (0, 0))


@pytest.mark.parametrize("z", [0.06, 0.075, 0.1, 0.15, 0.25])
@pytest.mark.parametrize("sign, obliquity", [(+1, False), (-1, True)])
def test_lattice_near_field_series(pkg, ctx, oracle, workloads, z, sign, obliquity, tiles):
    """Targets 0.06 - 0.25 m from a 10 um lattice, up to 8 mm off axis: the
    far-field GEMM's bounds fail for most (target, tile) pairs there.  The
    16-row kernels take the near-field GEMM (cross factor to second order,
    option grid_near), then the q-series to third order per pair (option
    grid_series), then exact pairs; each combination stays within 1e-6 of the
    oracle, and the near-field GEMM is a different path (other rounding)."""
    u = workloads.grid(256, 192, 10e-6, 0.0)
    g = np.random.default_rng(int(z * 1000))
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    v = targets(int(z * 1000), 900, z, half=8e-3)
    sub = np.arange(0, 900, 3)
    ref = oracle.transform_full(sign, LAM, amp, phase, u, v[sub], obliquity=obliquity)
    plan = ctx.plan(u)
    ys = {}
    for near, series in NEAR_PATHS:
        pkg.set_option("grid_near", near)
        pkg.set_option("grid_series", series)
        y = plan.transform(amp, phase, v, sign, LAM, pkg.FP32_SPLIT, obliquity=obliquity)
        err = rel_l2(y[sub], ref)
        print(f"z={z} near={near} series={series}: rel L2 {err:.3e}")
        assert err <= 1e-6
        ys[near, series] = y
    if tiles != 8 and z < 0.2:
        assert not np.array_equal(ys[1, 1], ys[0, 1])


@pytest.mark.parametrize("seed", range(8))
def test_lattice_random_near_field(pkg, ctx, seed, tiles):
    """Seeded near-field geometries (4 - 20 cm, targets up to 1 cm off axis,
    pitches 4 - 16 um, 400 - 700 nm, either sign, with or without obliquity),
    where the far-field GEMM bounds fail for most (target, tile) pairs and the
    near-field GEMM, the series path and the exact path share the work: within
    1e-6 of FP64 with each of them switched on and off."""
    g = np.random.default_rng(2000 + seed)
    nx, ny = 16 * int(g.integers(2, 17)), 32 * int(g.integers(1, 7))
    px, py = g.uniform(4e-6, 16e-6, 2)
    x = (np.arange(nx) - (nx - 1) / 2.0) * px
    y = (np.arange(ny) - (ny - 1) / 2.0) * py
    X, Y = np.meshgrid(x, y, indexing="ij")
    u = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    z = float(np.exp(g.uniform(np.log(0.04), np.log(0.2))))
    v = np.zeros((int(g.integers(300, 900)), 3))
    v[:, :2] = (g.random((len(v), 2)) - 0.5) * 2 * g.uniform(2e-3, 1e-2)
    v[:, 2] = z
    lam = g.uniform(400e-9, 700e-9)
    sign = 1 if g.random() < 0.5 else -1
    obl = bool(g.random() < 0.4)
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    plan = ctx.plan(u)
    assert plan.lattice() == (nx, ny)
    y64 = plan.transform(amp, phase, v, sign, lam, pkg.FP64, obliquity=obl)
    for near, series in NEAR_PATHS:
        pkg.set_option("grid_near", near)
        pkg.set_option("grid_series", series)
        y32 = plan.transform(amp, phase, v, sign, lam, pkg.FP32_SPLIT, obliquity=obl)
        err = rel_l2(y32, y64)
        print(f"seed {seed}: {nx}x{ny}, z {z:.3f} m, near {near} series {series}: "
              f"rel L2 {err:.3e}")
        assert err <= 1e-6


@pytest.mark.parametrize("z", [0.07, 0.12])
def test_lattice_near_field_complex_weights_and_adjoint(pkg, ctx, oracle, workloads, z, tiles):
    """Complex weights (the adjoint's sources) on a lattice plan with targets in
    the near field, and the adjoint of such a plane: the near-field GEMM takes
    the weights' real form like the far-field one."""
    u = workloads.grid(160, 128, 8e-6, 0.0)
    g = np.random.default_rng(int(z * 1000))
    gz = g.standard_normal(len(u)) + 1j * g.standard_normal(len(u))
    v = targets(int(z * 100), 700, z, half=6e-3)
    plan = ctx.plan(u)
    for near in (1, 0):
        pkg.set_option("grid_near", near)
        yc = plan.transform_complex(gz, v, -1, LAM, pkg.FP32_SPLIT)
        assert rel_l2(yc, oracle.transform_cartesian(-1, LAM, gz, u, v)) <= 1e-6
        amp, theta = weights(int(z * 1000) + 1, len(v))
        gt = g.standard_normal(len(u)) + 1j * g.standard_normal(len(u))
        ga, gth = plan.adjoint(gt, amp, theta, v, +1, LAM, pkg.FP32_SPLIT)
        ra, rt = oracle.adjoint(+1, LAM, amp, theta, v, gt, u)
        assert rel_l2(ga, ra) <= 1e-6 and rel_l2(gth, rt) <= 1e-6


def test_lattice_near_field_full_plane(pkg, ctx, oracle, workloads, tiles):
    """The benchmark's 1024^2 lattice with the 1024^2 target plane moved to
    z = 0.1 m: FP32-split within 1e-6 of the FP64 kernel over every target and
    of the oracle on a stride subset; the library's error report agrees."""
    import torch
    st = workloads.plane_1024(0).stages[0]
    v = st.v.copy()
    v[:, 2] = 0.1
    dev = [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in (st.amp, st.phase,
                                                                           st.u, v)]
    plan = ctx.plan(dev[2])
    y64 = plan.transform(dev[0], dev[1], dev[3], st.sign, st.lam, pkg.FP64).cpu().numpy()
    y32 = plan.transform(dev[0], dev[1], dev[3], st.sign, st.lam, pkg.FP32_SPLIT,
                         report_error=True).cpu().numpy()
    err_report = float(ctx.last_rel_l2.item())
    err = rel_l2(y32, y64)
    sub = workloads.stride_subset(len(v), 48)
    ref = oracle.transform_full(st.sign, st.lam, st.amp, st.phase, st.u, v, targets=sub)
    print(f"1024^2 at 0.1 m, tiles {tiles}: rel L2 {err:.3e} (report {err_report:.3e})")
    assert err <= 1e-6 and rel_l2(y32[sub], ref) <= 1e-6 and rel_l2(y64[sub], ref) <= 1e-12
    assert err_report <= 1e-6


@pytest.mark.parametrize("z", [2e-3, 2e-2])
@pytest.mark.parametrize("sign, obliquity", [(+1, False), (-1, True)])
def test_lattice_exact_rows_same_bits(pkg, ctx, oracle, workloads, z, sign, obliquity, tiles):
    """Targets 2 mm and 2 cm from the source plane: the (target, tile) pairs
    fail every expansion's bounds and are summed exactly.  The 16-row kernels
    take each pair's x and y from the lattice's row and column tables (option
    grid_rows); with the option off they gather the source's position instead.
    Both give the same bits: on a 40 x 72 lattice (16-row tiles of 8 rows and
    of 8 columns, and one tile column of padding only, since 5 tile columns
    become 6), on the generator's padded plane of 3001 points (54 x 55 and 31
    copies of point 0), with source splits, and within 1e-6 of the oracle."""
    planes = [workloads.grid(40, 72, 10e-6, 0.0),
              workloads.reference_plane(3001, 0.6e-3, 0.0, 1.0, (0.0, 0.0))]
    for u in planes:
        g = np.random.default_rng(len(u))
        amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
        v = targets(len(u) % 89, 320, z, half=1.5e-3)
        plan = ctx.plan(u)
        assert plan.lattice() in ((40, 72), (54, 55))
        ref = oracle.transform_full(sign, LAM, amp, phase, u, v, obliquity=obliquity)
        for splits in (0, 3):
            pkg.set_option("splits", splits)
            ys = []
            for rows in (1, 0):
                pkg.set_option("grid_rows", rows)
                ys.append(plan.transform(amp, phase, v, sign, LAM, pkg.FP32_SPLIT,
                                         obliquity=obliquity))
            pkg.set_option("grid_rows", 1)
            err = rel_l2(ys[0], ref)
            print(f"{len(u)} sources, z={z}, splits {splits}: rel L2 {err:.3e}")
            assert np.array_equal(ys[0], ys[1]) and err <= 1e-6
        pkg.set_option("splits", 0)
        # complex weights (the adjoint's sources) through the same exact pairs
        gz = g.standard_normal(len(u)) + 1j * g.standard_normal(len(u))
        yc = []
        for rows in (1, 0):
            pkg.set_option("grid_rows", rows)
            yc.append(plan.transform_complex(gz, v, sign, LAM, pkg.FP32_SPLIT))
        pkg.set_option("grid_rows", 1)
        refc = oracle.transform_cartesian(sign, LAM, gz, u, v)
        assert np.array_equal(yc[0], yc[1]) and rel_l2(yc[0], refc) <= 1e-6


@pytest.mark.parametrize("splits", [1, 3, 7])
def test_lattice_tail_of_arbitrary_points(pkg, ctx, oracle, workloads, splits, tiles):
    """A uniform lattice followed by fewer than max(nx, ny) other sources at
    any positions (here: off the plane, off the lattice, and one 2 mm from some
    targets), summed exactly by every target in the first source split; with
    source splits 1, 3 and 7 and targets near and far enough that some waves
    of the 16 x 32 kernel fall back to 16 x 16 tiles."""
    lat = workloads.grid(64, 96, 10e-6, 0.0)
    g = np.random.default_rng(splits)
    tail = np.zeros((40, 3))
    tail[:, :2] = (g.random((40, 2)) - 0.5) * 2e-3
    tail[:, 2] = (g.random(40) - 0.5) * 2e-4
    tail[0] = [1e-3, -1e-3, 0.248]  # 2 mm below the target plane
    u = np.concatenate([lat, tail])
    amp, phase = g.random(len(u)), g.random(len(u)) * 2 * np.pi
    v = np.concatenate([targets(7, 400, 0.25, half=2e-3), targets(8, 400, 0.25, half=6e-3)])
    plan = ctx.plan(u)
    assert plan.lattice() == (64, 96)
    pkg.set_option("splits", splits)
    y = plan.transform(amp, phase, v, -1, LAM, pkg.FP32_SPLIT)
    ref = oracle.transform_full(-1, LAM, amp, phase, u, v)
    assert rel_l2(y, ref) <= 1e-6
    # the tail is summed: without it the field differs
    amp2 = amp.copy()
    amp2[len(lat):] = 0.0
    y2 = plan.transform(amp2, phase, v, -1, LAM, pkg.FP32_SPLIT)
    assert rel_l2(y2, y) > 1e-3
