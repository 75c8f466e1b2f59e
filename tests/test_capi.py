"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/holo.h declares, and its host-only entry points behave (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "holo.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(holo_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("holo_transform_full", "holo_transform_mc", "holo_plan_adjoint",
                 "holo_normalize_polar", "holo_create", "holo_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version_and_status_strings(pkg):
    L = pkg.lib()
    assert L.holo_abi_version() == 2
    assert L.holo_status_string(0) == b"ok"
    assert L.holo_status_string(1) == b"invalid argument"
    assert L.holo_status_string(99) == b"unknown status"


def test_options(pkg):
    pkg.set_option("r_f64", 4)
    assert pkg.get_option("r_f64") == 4
    with pytest.raises(pkg.HoloError):
        pkg.set_option("r_f64", 3)
    with pytest.raises(pkg.HoloError):
        pkg.set_option("no_such_option", 1)
    assert pkg.get_option("no_such_option") == -1
    assert pkg.get_option("mc_passes") == 0
    pkg.set_option("mc_passes", 3)
    assert pkg.get_option("mc_passes") == 3
    with pytest.raises(pkg.HoloError):
        pkg.set_option("mc_passes", 65)
    pkg.set_option("mc_passes", 0)
    for name in ("f64_libm", "f64_planar", "f64_rows", "grid", "grid16", "grid32", "grid64", "grid_series",
                 "grid_near", "grid_rows", "mc_lattice"):
        old = pkg.get_option(name)
        assert old in (0, 1)
        pkg.set_option(name, 1 - old)
        assert pkg.get_option(name) == 1 - old
        pkg.set_option(name, old)


def _write_run(pkg, d, len_y=5):
    """A run directory as the CLI writes it (cuda/util.h:358-444): one x and
    one y record."""
    import json
    g = np.random.default_rng(1)
    rec = []
    for k, pos, n in (("x0", "u0", 3), ("y0", "v0", len_y)):
        pkg.write_dat(str(d / f"{k}_amp.dat"), g.random(n))
        pkg.write_dat(str(d / f"{k}_phase.dat"), g.random(n))
        pkg.write_dat(str(d / f"{pos}.dat"), g.random(3 * n))
        rec.append({"phasor": k, "amp_sum": 1.0, "phase_sum": 0.0, "pos": pos, "len": n,
                    "precision": 8, "dims": 3, "x_offset": 0, "y_offset": 0, "z_offset": 0,
                    "width": 1e-3, "randomized": 0, "aspect_ratio": 1, "timestamp": 0,
                    "runtime": 0.0, "flops": 0.0})
    with open(d / "out.json", "w") as f:
        for r in rec:
            f.write(json.dumps(r) + "\n")


def test_parse_file_directory_and_zip(pkg, tmp_path):
    """py/util.py:700-744: the reference's reader layout, from the files and
    from out.zip."""
    import zipfile
    _write_run(pkg, tmp_path)
    params, data = pkg.parse_file(str(tmp_path), zipfilename=None)
    assert [p["phasor"] for p in params["x"] + params["y"]] == ["x0", "y0"]
    assert data["x"][0].shape == (3, 2) and data["y"][0].shape == (5, 2)
    assert data["u"][0].shape == (3, 3) and data["v"][0].shape == (5, 3)
    assert np.array_equal(data["y"][0][:, 0], pkg.read_dat(str(tmp_path / "y0_amp.dat")))
    assert np.array_equal(data["v"][0].reshape(-1), pkg.read_dat(str(tmp_path / "v0.dat")))
    with zipfile.ZipFile(tmp_path / "out.zip", "w") as z:
        for name in os.listdir(tmp_path):
            if name != "out.zip":
                z.write(tmp_path / name, name)
    for name in os.listdir(tmp_path):
        if name != "out.zip":
            os.remove(tmp_path / name)
    params2, data2 = pkg.parse_file(str(tmp_path))
    assert params2 == params
    for k in "xyuv":
        assert np.array_equal(data2[k][0], data[k][0])
    assert pkg.parse_file(str(tmp_path), read_data=False) == (params2,)


def test_parse_file_short_record(pkg, tmp_path):
    """A record whose files are shorter than its len gets len 0 (py/util.py:738-741)."""
    _write_run(pkg, tmp_path)
    pkg.write_dat(str(tmp_path / "y0_amp.dat"), np.zeros(2))
    params, data = pkg.parse_file(str(tmp_path), zipfilename=None)
    assert params["y"][0]["len"] == 0 and data["y"] == [] and data["v"] == []
    assert len(data["x"]) == 1


def test_dat_roundtrip_is_raw_little_endian_f8(pkg, tmp_path):
    """cuda/util.h:248-337: headerless float64; py/util.py:668-678 reads '<f8'."""
    x = np.random.default_rng(0).standard_normal(1001)
    p = str(tmp_path / "x0_amp.dat")
    pkg.write_dat(p, x)
    assert os.path.getsize(p) == 8 * x.size
    assert np.array_equal(np.fromfile(p, dtype="<f8"), x)
    assert np.array_equal(pkg.read_dat(p), x)


def test_dat_errors(pkg, tmp_path):
    empty = tmp_path / "empty.dat"
    empty.write_bytes(b"")
    ragged = tmp_path / "ragged.dat"
    ragged.write_bytes(b"\x00" * 12)
    for p in (str(empty), str(ragged), str(tmp_path / "missing.dat")):
        with pytest.raises(pkg.HoloError) as e:
            pkg.read_dat(p)
        assert e.value.status == 4
    assert "missing.dat" in pkg.lib().holo_last_error().decode() or True


def test_create_without_device_fails_cleanly(pkg):
    """With no gfx950 device the context refuses with HOLO_EDEVICE (no exit)."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert pkg.lib().holo_create(ctypes.byref(h), 0, None) == 3
    assert pkg.lib().holo_last_error()
    with pytest.raises(pkg.HoloError):
        pkg.Context(0)


def test_null_and_bad_arguments_need_no_device(pkg):
    L = pkg.lib()
    assert L.holo_transform_full(None, 0, 1, 1e-6, None, None, None, 0, None, 0, None, None,
                                 0) == 1
    assert L.holo_plan_info(None, None, None, None, None) == 1
    assert L.holo_set_option(None, 0) == 1
