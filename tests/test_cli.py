"""The `holo` command line (flags and files of cuda/main.cu + cuda/input.h)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "holographic-projector-simulations_amd", "bin", "holo")
GOLDEN = os.path.join(ROOT, "tests", "golden")
LAM_REF = 0.65e-6  # the CLI's default wavelength (cuda/macros.h:23-25)


def run(args, cwd=None, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, env=e, timeout=120)


@pytest.fixture(scope="module", autouse=True)
def _built():
    # a missing binary is a build failure, not a reason to skip the boundary tests
    assert os.path.exists(CLI), f"{CLI} is missing: run __graft_entry__.build()"


def test_help_exits_zero():
    r = run(["-h"])
    assert r.returncode == 0 and "Usage" in r.stdout and "-f {dir}" in r.stdout


@pytest.mark.parametrize("args, msg", [
    (["-p", "4"], "algorithm"),
    (["-p", "0"], "algorithm"),
    (["-y", "0"], "-y"),
    (["-b", "64", "-B", "32"], "blockDim"),
    (["-w", "0"], "z offset"),
    (["-w", "0.1", "-m", "-0.1"], "projection plane"),
])
def test_invalid_arguments_exit_nonzero(args, msg):
    """cuda/input.h:129-149 asserts become messages and exit status 1."""
    r = run(args)
    assert r.returncode == 1 and msg in r.stderr


def test_missing_and_ragged_inputs(tmp_path):
    r = run(["-f", str(tmp_path)])
    assert r.returncode == 1 and "x0_amp.dat" in r.stderr
    for name in ("x0_amp.dat", "x0_phase.dat", "u0.dat", "v0.dat"):
        shutil.copy(os.path.join(GOLDEN, "plain_small_" + {"x0_amp.dat": "x_amp.dat",
                                                         "x0_phase.dat": "x_phase.dat",
                                                         "u0.dat": "u.dat",
                                                         "v0.dat": "v.dat"}[name]),
                    tmp_path / name)
    (tmp_path / "u0.dat").write_bytes((tmp_path / "u0.dat").read_bytes()[:-24])
    r = run(["-f", str(tmp_path)])
    assert r.returncode == 1 and "source counts" in r.stderr


def test_invalid_obliquity_env_exits_nonzero():
    r = run(["-h"], env={"HOLO_OBLIQUITY": "2"})
    assert r.returncode == 0  # -h prints usage before the environment is read
    r = run(["-x", "1"], env={"HOLO_OBLIQUITY": "yes"})
    assert r.returncode == 1 and "HOLO_OBLIQUITY" in r.stderr


def stage_golden(d):
    for src, dst in (("x_amp", "x0_amp"), ("x_phase", "x0_phase"), ("u", "u0"), ("v", "v0")):
        shutil.copy(os.path.join(GOLDEN, f"plain_small_{src}.dat"), os.path.join(d, f"{dst}.dat"))


@pytest.mark.gpu
@pytest.mark.parametrize("projection", [False, True])
def test_file_mode_matches_oracle(tmp_path, oracle, projection):
    """holo -f DIR: y = time_transform<Backwards>(x, u, v) written as y0_*.dat
    with the reference's normalisation and polar conversion; -F also writes the
    forwards transform as z0_0 (cuda/main.cu:203-222)."""
    stage_golden(tmp_path)
    amp = np.fromfile(tmp_path / "x0_amp.dat", "<f8")
    phase = np.fromfile(tmp_path / "x0_phase.dat", "<f8")
    u = np.fromfile(tmp_path / "u0.dat", "<f8")
    v = np.fromfile(tmp_path / "v0.dat", "<f8")
    M = v.size // 3
    if projection:
        shutil.copy(tmp_path / "v0.dat", tmp_path / "w0.dat")
    args = ["-f", str(tmp_path), "-y", str(M)] + (["-F"] if projection else [])
    r = run(args)
    assert r.returncode == 0, r.stderr
    ya = np.fromfile(tmp_path / "y0_amp.dat", "<f8")
    yp = np.fromfile(tmp_path / "y0_phase.dat", "<f8")
    ref = oracle.transform_full(-1, LAM_REF, amp, phase, u, v)
    ra, rp, _ = oracle.normalize_polar(ref)
    assert ya.size == M and np.max(np.abs(ya - ra)) < 1e-12
    z = ya * np.exp(1j * yp)
    assert np.linalg.norm(z - ra * np.exp(1j * rp)) / np.linalg.norm(ra) < 1e-12
    assert np.array_equal(np.fromfile(tmp_path / "v0.dat", "<f8"), v)
    lines = [json.loads(l) for l in open(tmp_path / "out.json")]
    assert lines[0]["phasor"] == "y0" and lines[0]["pos"] == "v0" and lines[0]["len"] == M
    assert lines[0]["precision"] == 8 and lines[0]["dims"] == 3
    if projection:
        za = np.fromfile(tmp_path / "z0_0_amp.dat", "<f8")
        rf = oracle.transform_full(+1, LAM_REF, amp, phase, u, v)
        fa, _, _ = oracle.normalize_polar(rf)
        assert np.max(np.abs(za - fa)) < 1e-12
        assert lines[1]["phasor"] == "z0_0" and lines[1]["pos"] == "w0_0"


@pytest.mark.gpu
def test_target_count_comes_from_y_flag(tmp_path, oracle):
    """M is -y, not the size of v0.dat: extra targets are the origin."""
    stage_golden(tmp_path)
    v = np.fromfile(tmp_path / "v0.dat", "<f8")
    M = v.size // 3
    r = run(["-f", str(tmp_path), "-y", str(M + 10)])
    assert r.returncode == 0, r.stderr
    ya = np.fromfile(tmp_path / "y0_amp.dat", "<f8")
    assert ya.size == M + 10
    amp = np.fromfile(tmp_path / "x0_amp.dat", "<f8")
    phase = np.fromfile(tmp_path / "x0_phase.dat", "<f8")
    u = np.fromfile(tmp_path / "u0.dat", "<f8")
    vv = np.concatenate([v, np.zeros(30)])
    ra, _, _ = oracle.normalize_polar(oracle.transform_full(-1, LAM_REF, amp, phase, u, vv))
    assert np.max(np.abs(ya - ra)) < 1e-12


@pytest.mark.gpu
def test_generated_mode_writes_reference_files(tmp_path, pkg):
    """No -f: object circle -> projector lattice -> projection lattice in ../tmp/,
    readable by the reference's reader layout (parse_file)."""
    work = tmp_path / "run"
    (tmp_path / "tmp").mkdir()
    work.mkdir()
    r = run(["-x", "16", "-y", "4096", "-z", "1024", "-Z", "2"], cwd=str(work))
    assert r.returncode == 0, r.stderr
    out = tmp_path / "tmp"
    lines = [json.loads(l) for l in open(out / "out.json")]
    assert [l["phasor"] for l in lines] == ["x0", "y0", "z0_0", "z0_1"]
    y = np.fromfile(out / "y0_amp.dat", "<f8")
    assert y.size == 4096 and abs(y.max() - 1.0) < 1e-12
    w = np.fromfile(out / "w0_1.dat", "<f8").reshape(-1, 3)
    assert w.shape == (1024, 3) and np.allclose(w[:, 2], 0.35)
    u = np.fromfile(out / "u0.dat", "<f8").reshape(-1, 3)
    assert np.allclose(np.hypot(u[:, 0], u[:, 1]), 0.5e-3)  # circle of width 1 mm
    params, data = pkg.parse_file(str(out), zipfilename=None)
    assert [len(params[k]) for k in "xyz"] == [1, 1, 2]
    assert data["x"][0].shape == (16, 2) and data["u"][0].shape == (16, 3)
    assert data["y"][0].shape == (4096, 2) and np.array_equal(data["y"][0][:, 0], y)
    assert [a.shape for a in data["z"]] == [(1024, 2)] * 2
    assert np.array_equal(data["w"][1], w)


@pytest.mark.gpu
def test_generated_plane_padding(tmp_path, workloads):
    """A projector point count that is not x * y (-y 1000): a 31 x 32 lattice
    in x-slow order followed by 8 copies of point 0 (cuda/init.h:96-157), the
    same layout as workloads.reference_plane; the field over it is finite."""
    work = tmp_path / "run"
    (tmp_path / "tmp").mkdir()
    work.mkdir()
    r = run(["-x", "16", "-y", "1000", "-z", "1024"], cwd=str(work))
    assert r.returncode == 0, r.stderr
    v = np.fromfile(tmp_path / "tmp" / "v0.dat", "<f8").reshape(-1, 3)
    assert v.shape == (1000, 3)
    lat = v[:992].reshape(31, 32, 3)
    assert np.all(lat[:, :, 0] == lat[:, :1, 0]) and np.all(lat[:, :, 1] == lat[:1, :, 1])
    assert np.all(np.diff(lat[:, 0, 0]) > 0) and np.all(np.diff(lat[0, :, 1]) > 0)
    assert np.array_equal(v[992:], np.repeat(v[:1], 8, 0))
    ref = workloads.reference_plane(1000, 1.0, 0.0)
    dx = np.diff(lat[:, 0, 0]).mean() / np.diff(ref[:992:32, 0]).mean()
    np.testing.assert_allclose(v[:, :2] - v[:1, :2], (ref[:, :2] - ref[:1, :2]) * dx,
                               rtol=0, atol=1e-12 * dx)
    y = np.fromfile(tmp_path / "tmp" / "y0_amp.dat", "<f8")
    assert y.size == 1000 and np.all(np.isfinite(y))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["circle", "line", "logline", "dottedcircle"])
def test_generated_object_shapes(tmp_path, shape):
    """HOLO_SHAPE selects the object of cuda/init.h:160-234 (width 1 mm, 16
    points, first plane: no lateral offset, modulate 0), checked against the
    closed forms."""
    work = tmp_path / "run"
    (tmp_path / "tmp").mkdir()
    work.mkdir()
    r = run(["-x", "16", "-y", "1024", "-z", "1024"], cwd=str(work), env={"HOLO_SHAPE": shape})
    assert r.returncode == 0, r.stderr
    u = np.fromfile(tmp_path / "tmp" / "u0.dat", "<f8").reshape(-1, 3)
    n, w = 16, 1e-3
    i = np.arange(n)
    if shape == "line":
        x, y = i * (w / (n - 1)) - w / 2, np.zeros(n)
    elif shape == "logline":
        x, y = 10.0 ** (i * (np.log10(1 + w) / (n - 1))) - 1 + w / 2, np.zeros(n)
    else:
        a = 2 * np.pi * i / (n if shape == "circle" else n - 1) + 0.1125
        x, y = np.sin(a) * w / 2, np.cos(a) * w / 2
        if shape == "dottedcircle":
            x[0] = y[0] = 0.0
    assert u.shape == (n, 3)
    np.testing.assert_allclose(u[:, 0], x, rtol=0, atol=1e-15)
    np.testing.assert_allclose(u[:, 1], y, rtol=0, atol=1e-15)
    assert np.all(u[:, 2] == u[0, 2])


def test_invalid_shape_exits_nonzero():
    r = run(["-x", "4"], env={"HOLO_SHAPE": "square"})
    assert r.returncode != 0 and "HOLO_SHAPE" in r.stderr


DROPIN = os.path.join(ROOT, "holographic-projector-simulations_amd", "_build", "transform_dropin")


@pytest.mark.gpu
def test_cpp_dropin_templates(tmp_path, oracle):
    """include/holo/transform.hpp used as cuda/main.cu uses the reference's
    templates (tests/cpp/transform_dropin.cpp): transform_full<Backwards,
    Reduced, true> returns the padded field of the oracle's sum, time_transform
    its normalised polar form, and a target in the first source's plane is
    refused as the reference asserts (cuda/transform.cu:193)."""
    assert os.path.exists(DROPIN), f"{DROPIN} is missing: run __graft_entry__.build()"
    r = subprocess.run([DROPIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "caught EINVAL" in r.stdout
    rd = lambda n: np.fromfile(tmp_path / n, "<f8")
    amp, phase, u, v = rd("x_amp.dat"), rd("x_phase.dat"), rd("u.dat"), rd("v.dat")
    M = v.size // 3
    y = rd("y.dat").view(np.complex128)
    ref = oracle.transform_full(-1, LAM_REF, amp, phase, u, v)
    assert y.size == 2048 and np.all(y[M:] != 0)  # padded targets sit at the origin
    assert np.linalg.norm(y[:M] - ref) / np.linalg.norm(ref) <= 1e-12
    # the padded targets enter the max-normalisation, as in the reference
    pad = np.zeros((y.size, 3))
    pad[:M] = v.reshape(-1, 3)
    full = oracle.transform_full(-1, LAM_REF, amp, phase, u, pad.reshape(-1))
    ra, rp, _ = oracle.normalize_polar(full)
    pa, pp = rd("p_amp.dat"), rd("p_phase.dat")
    assert pa.size == M
    assert np.max(np.abs(pa - ra[:M])) <= 1e-12
    assert np.max(np.abs(np.angle(np.exp(1j * (pp - rp[:M]))))) <= 1e-9


@pytest.mark.gpu
def test_reference_wave_mode(tmp_path, oracle):
    """HOLO_PROJECT_PHASE=0 (the reference's PROJECT_PHASE=0 build,
    cuda/transform.cu:379-395): y is normalised, a unit point source 0.25 m
    behind the v plane adds a normalised reference wave, and the projector
    keeps |y + y_ref| with the reference wave's phase."""
    stage_golden(tmp_path)
    v = np.fromfile(tmp_path / "v0.dat", "<f8")
    M = v.size // 3
    r = run(["-f", str(tmp_path), "-y", str(M)], env={"HOLO_PROJECT_PHASE": "0"})
    assert r.returncode == 0, r.stderr
    ya = np.fromfile(tmp_path / "y0_amp.dat", "<f8")
    yp = np.fromfile(tmp_path / "y0_phase.dat", "<f8")
    amp = np.fromfile(tmp_path / "x0_amp.dat", "<f8")
    phase = np.fromfile(tmp_path / "x0_phase.dat", "<f8")
    u = np.fromfile(tmp_path / "u0.dat", "<f8")
    y = oracle.transform_full(-1, LAM_REF, amp, phase, u, v)
    y = y / np.abs(y).max()
    ref = oracle.transform_full(+1, LAM_REF, [1.0], [0.0], [0.0, 0.0, v[2] - 0.25], v)
    ref = ref / np.abs(ref).max()
    assert np.max(np.abs(ya - np.abs(y + ref))) < 1e-12
    assert np.max(np.abs(np.angle(np.exp(1j * (yp - np.angle(ref)))))) < 1e-9


@pytest.mark.gpu
def test_file_mode_obliquity(tmp_path, oracle):
    """HOLO_OBLIQUITY=1: every pair also carries |dz| / d (the MATLAB model's
    cosine factor, matlab/holofunctions.m:226-227)."""
    stage_golden(tmp_path)
    v = np.fromfile(tmp_path / "v0.dat", "<f8")
    M = v.size // 3
    r = run(["-f", str(tmp_path), "-y", str(M)], env={"HOLO_OBLIQUITY": "1"})
    assert r.returncode == 0, r.stderr
    ya = np.fromfile(tmp_path / "y0_amp.dat", "<f8")
    yp = np.fromfile(tmp_path / "y0_phase.dat", "<f8")
    amp = np.fromfile(tmp_path / "x0_amp.dat", "<f8")
    phase = np.fromfile(tmp_path / "x0_phase.dat", "<f8")
    u = np.fromfile(tmp_path / "u0.dat", "<f8")
    ra, rp, _ = oracle.normalize_polar(oracle.transform_full(-1, LAM_REF, amp, phase, u, v,
                                                             obliquity=True))
    assert np.max(np.abs(ya - ra)) < 1e-12
    assert np.linalg.norm(ya * np.exp(1j * yp) - ra * np.exp(1j * rp)) / np.linalg.norm(ra) < 1e-12


@pytest.mark.gpu
def test_matlab_projector_fixture(tmp_path):
    """The input set of matlab/projector.m (32x32 version, tests/golden/
    matlab_projector_*) through `holo -f DIR -F`: y0 and z0_0 equal the
    committed expected outputs."""
    for k in ("x0_amp", "x0_phase", "u0", "v0", "w0"):
        shutil.copy(os.path.join(GOLDEN, f"matlab_projector_{k}.dat"), tmp_path / f"{k}.dat")
    r = run(["-f", str(tmp_path), "-y", "1024", "-F"])
    assert r.returncode == 0, r.stderr
    for k in ("y0", "z0_0"):
        a = np.fromfile(tmp_path / f"{k}_amp.dat", "<f8")
        p = np.fromfile(tmp_path / f"{k}_phase.dat", "<f8")
        ea = np.fromfile(os.path.join(GOLDEN, f"matlab_projector_{k}_amp.dat"), "<f8")
        ep = np.fromfile(os.path.join(GOLDEN, f"matlab_projector_{k}_phase.dat"), "<f8")
        assert np.max(np.abs(a - ea)) < 1e-12
        assert np.linalg.norm(a * np.exp(1j * p) - ea * np.exp(1j * ep)) / np.linalg.norm(ea) < 1e-12


def generated_run(tmp_path, args, env=None):
    """Run the generated mode in tmp_path/run (it writes to ../tmp/); return
    the output directory and the parsed out.json records by phasor name."""
    work = tmp_path / "run"
    (tmp_path / "tmp").mkdir(parents=True)
    work.mkdir()
    r = run(args, cwd=str(work), env=env)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "tmp"
    recs = {l["phasor"]: l for l in (json.loads(x) for x in open(out / "out.json"))}
    return out, recs


def read_polar(out, k):
    a = np.fromfile(out / f"{k}_amp.dat", "<f8")
    p = np.fromfile(out / f"{k}_phase.dat", "<f8")
    return a, p


def lattice_sites(n, width, z, aspect, offset):
    """The reference generator's unjittered sites (cuda/init.h:96-157) and pitches."""
    x = int(np.sqrt(n * aspect))
    y = n // x
    height = width / aspect
    dx, dy = width / x, height / y
    i, j = np.meshgrid(np.arange(x), np.arange(y), indexing="ij")
    s = np.zeros((x * y, 3))
    s[:, 0] = (i * dx - (width - dx) / 2 + offset[0]).ravel()
    s[:, 1] = (j * dy - (height - dy) / 2 + offset[1]).ravel()
    s[:, 2] = z
    return s, dx, dy


@pytest.mark.gpu
def test_randomized_projector_pixels(tmp_path, oracle, pkg):
    """-r (cuda/init.h:116-147): every pixel of the projector and projection
    lattices moves by up to (1 - 0.025) pitch / 2 per axis from its site, the
    points left over copy the (jittered) point 0, the jittered plane is no
    lattice, and the fields over the written positions match the oracle (FP64
    <= 1e-12; FP32-split <= 1e-6 on the tiled kernel, with the error report
    recorded in out.json)."""
    n_y, n_z = 4100, 1024  # 64 x 64 + 4 copies of point 0
    out, recs = generated_run(tmp_path, ["-x", "16", "-y", str(n_y), "-z", str(n_z), "-r"])
    v = np.fromfile(out / "v0.dat", "<f8").reshape(-1, 3)
    w = np.fromfile(out / "w0_0.dat", "<f8").reshape(-1, 3)
    for pts, n, width in ((v, n_y, 512 * 7e-6), (w, n_z, 1.2e-3)):
        z = pts[0, 2]  # 0, and the object plane's z = 10^log10(0.35) for w
        sites, dx, dy = lattice_sites(n, width, z, 1.0, (0.0, 0.0))
        body = len(sites)
        d = pts[:body, :2] - sites[:, :2]
        assert np.all(np.abs(d[:, 0]) <= 0.4875 * dx * (1 + 1e-12))
        assert np.all(np.abs(d[:, 1]) <= 0.4875 * dy * (1 + 1e-12))
        assert np.abs(d[:, 0]).max() > 0.4 * dx and np.abs(d[:, 1]).max() > 0.4 * dy
        # roughly uniform offsets: mean near 0, spread near that of U(-0.4875, 0.4875)
        assert abs(d[:, 0].mean() / dx) < 0.02 and abs(d[:, 0].std() / dx - 0.975 / 12 ** 0.5) < 0.02
        assert np.all(pts[:, 2] == z)
        assert np.array_equal(pts[body:], np.repeat(pts[:1], n - body, 0))
        assert pkg.lattice_layout(pts)[:2] == (0, 0)
    assert abs(w[0, 2] - 0.35) <= 1e-15
    assert recs["y0"]["randomized"] == 1 and recs["z0_0"]["randomized"] == 1
    u = np.fromfile(out / "u0.dat", "<f8")
    xa, xp = read_polar(out, "x0")
    ya, yp = read_polar(out, "y0")
    ry = oracle.transform_full(-1, LAM_REF, xa, xp, u, v)
    ra, rp, _ = oracle.normalize_polar(ry)
    assert np.max(np.abs(ya - ra)) < 1e-12
    assert np.linalg.norm(ya * np.exp(1j * yp) - ra * np.exp(1j * rp)) / np.linalg.norm(ra) < 1e-12
    za, zp = read_polar(out, "z0_0")
    rz = oracle.transform_full(+1, LAM_REF, ya, yp, v, w)
    rza, rzp, _ = oracle.normalize_polar(rz)
    assert np.linalg.norm(za * np.exp(1j * zp) - rza * np.exp(1j * rzp)) / np.linalg.norm(rza) < 1e-12
    assert "rel_l2_fp64" not in recs["z0_0"]
    # FP32-split over the same jittered projector (its own run: the jitter
    # stream restarts with the process, so the positions are the same)
    out2, recs2 = generated_run(tmp_path / "fp32", ["-x", "16", "-y", str(n_y), "-z", str(n_z),
                                                    "-r"], env={"HOLO_PRECISION": "fp32split"})
    assert np.array_equal(np.fromfile(out2 / "v0.dat", "<f8").reshape(-1, 3), v)
    ya2, yp2 = read_polar(out2, "y0")
    za2, zp2 = read_polar(out2, "z0_0")
    z2 = za2 * np.exp(1j * zp2)
    rz2 = oracle.transform_full(+1, LAM_REF, ya2, yp2, v, w)
    rza2, rzp2, _ = oracle.normalize_polar(rz2)
    err = np.linalg.norm(z2 - rza2 * np.exp(1j * rzp2)) / np.linalg.norm(rza2)
    assert err <= 1e-6
    assert 0 < recs2["z0_0"]["rel_l2_fp64"] <= 1e-6


@pytest.mark.gpu
def test_generated_planes_interpolated(tmp_path, oracle, workloads):
    """-X 2 -Z 2 with ranges (cuda/main.cu:125-199): object plane i sits at
    z = gerp(0.35 -> 0.25, i) (geometric interpolation, cuda/algebra.h:140-150)
    with lerp'd width and offsets; projection plane j at z offset
    lerp(0 -> 0.1, j) from it, width gerp(1.2 -> 2.4 mm, j), shifted into one
    quadrant by -q.  Every written position set is the generator's, every
    out.json record carries the plane's offsets and width, and every y<i> and
    z<i>_<j> is the oracle's normalised field over the written inputs."""
    args = ["-X", "2", "-Z", "2", "-x", "12", "-y", "1024", "-z", "256",
            "-w", "0.35", "-W", "0.25", "-u", "1e-4", "-U", "3e-4", "-v", "-2e-4", "-V", "0",
            "-o", "1e-3", "-O", "2e-3", "-n", "1.2e-3", "-N", "2.4e-3", "-m", "0", "-M", "0.1",
            "-q"]
    out, recs = generated_run(tmp_path, args)
    assert sorted(recs) == ["x0", "x1", "y0", "y1", "z0_0", "z0_1", "z1_0", "z1_1"]
    rec = lambda r, k: pytest.approx(r, rel=1e-5, abs=1e-12)  # out.json keeps 6 digits
    gerp = lambda a, b, r: 10 ** ((1 - r) * np.log10(a) + r * np.log10(b))
    for i in range(2):
        zo = gerp(0.35, 0.25, i)
        xo, yo = 1e-4 + i * 2e-4, -2e-4 + i * 2e-4
        xr = recs[f"x{i}"]
        assert xr["z_offset"] == rec(zo, 0) and xr["x_offset"] == rec(xo, 0)
        assert xr["y_offset"] == rec(yo, 0) and xr["width"] == rec(1e-3 * (1 + i), 0)
        u = np.fromfile(out / f"u{i}.dat", "<f8").reshape(-1, 3)
        assert np.allclose(np.hypot(u[:, 0] - xo, u[:, 1] - yo), 0.5e-3 * (1 + i), rtol=0,
                           atol=1e-15)
        assert np.all(u[:, 2] == u[0, 2]) and u[0, 2] == pytest.approx(zo, rel=1e-15)
        v = np.fromfile(out / f"v{i}.dat", "<f8").reshape(-1, 3)
        np.testing.assert_allclose(v, workloads.reference_plane(1024, 512 * 7e-6, 0.0),
                                   rtol=0, atol=1e-16)
        xa, xp = read_polar(out, f"x{i}")
        assert np.allclose(xa, np.linspace(1, 0, 12)) and np.all(xp == 0)
        ya, yp = read_polar(out, f"y{i}")
        ra, rp, _ = oracle.normalize_polar(oracle.transform_full(-1, LAM_REF, xa, xp, u, v))
        assert np.max(np.abs(ya - ra)) < 1e-12
        assert np.linalg.norm(ya * np.exp(1j * yp) - ra * np.exp(1j * rp)) / np.linalg.norm(ra) < 1e-12
        for j in range(2):
            width = gerp(1.2e-3, 2.4e-3, j)
            zr = recs[f"z{i}_{j}"]
            xz, yz, zz = xo + width / 2, yo + width / 2, zo + 0.1 * j
            assert zr["width"] == rec(width, 0) and zr["x_offset"] == rec(xz, 0)
            assert zr["y_offset"] == rec(yz, 0) and zr["z_offset"] == rec(zz, 0)
            w = np.fromfile(out / f"w{i}_{j}.dat", "<f8").reshape(-1, 3)
            np.testing.assert_allclose(
                w, workloads.reference_plane(256, width, zz, 1.0, (xz, yz)), rtol=0, atol=1e-15)
            za, zp = read_polar(out, f"z{i}_{j}")
            rza, rzp, _ = oracle.normalize_polar(oracle.transform_full(+1, LAM_REF, ya, yp, v, w))
            assert np.max(np.abs(za - rza)) < 1e-12
            assert (np.linalg.norm(za * np.exp(1j * zp) - rza * np.exp(1j * rzp))
                    / np.linalg.norm(rza)) < 1e-12
