import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

holo = __graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libholo.so")


@pytest.fixture(scope="session")
def pkg():
    return holo


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def workloads():
    import importlib
    return importlib.import_module("holo_amd.workloads")


@pytest.fixture(scope="session")
def ctx():
    c = holo.Context(0)
    yield c
    c.close()


@pytest.fixture(autouse=True)
def _reset_options():
    """Tuning knobs are process-wide: restore defaults around every test."""
    yield
    if os.path.exists(holo.LIB_PATH):
        for k, v in (("r_f64", 4), ("r_f32", 2), ("splits", 0), ("target_waves", 16384),
                     ("f64_libm", 0), ("f64_planar", 1), ("f64_rows", 1),
                     ("grid", 1), ("grid16", 1),
                     ("grid32", 1), ("grid64", 1), ("grid_series", 1), ("grid_near", 1), ("grid_rows", 1), ("mc_lattice", 1),
                     ("mc_passes", 0)):
            holo.set_option(k, v)
