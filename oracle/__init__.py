"""CPU oracle for the Huygens superposition -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this package, and only as the checker.  The
product library (``libholo.so``) never links, loads or calls it.

``holo_oracle.c`` restates the reference's FP64 formulation
(cuda/superposition.cu:35-51, cuda/kernel.cu:148-166, cuda/macros.h:23-31)
with a pinned operation order; ``numpy_baseline.py`` restates the
reference's NumPy CPU path (py/util.py:92-147).  Pinning of every function
is listed in the C file's header and exercised by ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_wavenumber.restype = ctypes.c_double
        L.oracle_wavenumber.argtypes = [ctypes.c_double]
        L.oracle_pair.restype = None
        L.oracle_pair.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_transform_full.restype = None
        L.oracle_transform_full.argtypes = [
            ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.c_size_t,
            _dp, ctypes.c_size_t, _i64p, ctypes.c_size_t, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_transform_cartesian.restype = None
        L.oracle_transform_cartesian.argtypes = [
            ctypes.c_int, ctypes.c_double, _dp, _dp, ctypes.c_size_t,
            _dp, ctypes.c_size_t, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_normalize_polar.restype = ctypes.c_double
        L.oracle_normalize_polar.argtypes = [_dp, ctypes.c_size_t, _dp, _dp]
        L.oracle_philox4x32_10.restype = None
        L.oracle_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.oracle_mc_index.restype = ctypes.c_uint64
        L.oracle_mc_index.argtypes = [ctypes.c_uint64] * 4
        L.oracle_mc_shared_index.restype = ctypes.c_uint64
        L.oracle_mc_shared_index.argtypes = [ctypes.c_uint64] * 4
        L.oracle_transform_mc.restype = None
        L.oracle_transform_mc.argtypes = [
            ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.c_size_t,
            _dp, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
            _i64p, ctypes.c_size_t, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_adjoint.restype = None
        L.oracle_adjoint.argtypes = [
            ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.c_size_t,
            _dp, _dp, ctypes.c_size_t, _dp, _dp, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def wavenumber(lam: float) -> float:
    return lib().oracle_wavenumber(lam)


def pair(sign: int, lam: float, a: float, phi: float, u, v, obliquity: bool = False):
    """One pair -> (re, im, d, arg) with the pinned operation order."""
    u = _f64(u); v = _f64(v)
    p = [ctypes.pointer(ctypes.c_double(0)) for _ in range(4)]
    lib().oracle_pair(sign, wavenumber(lam), a, phi, _d(u), _d(v), int(obliquity), *p)
    return tuple(x.contents.value for x in p)


def transform_full(sign: int, lam: float, amp, phase, u, v, targets=None,
                   nthreads: int = 0, obliquity: bool = False) -> np.ndarray:
    """Brute-force sum; returns complex128 (len(targets) or M)."""
    amp = _f64(amp); phase = _f64(phase); u = _f64(u).reshape(-1); v = _f64(v).reshape(-1)
    N, M = amp.size, v.size // 3
    assert phase.size == N and u.size == 3 * N
    if targets is not None:
        t = np.ascontiguousarray(targets, dtype=np.int64)
        assert t.size == 0 or (t.min() >= 0 and t.max() < M)
        cnt, tp = t.size, t.ctypes.data_as(_i64p)
    else:
        cnt, tp = M, None
    y = np.zeros(2 * cnt, dtype=np.float64)
    lib().oracle_transform_full(sign, lam, _d(amp), _d(phase), _d(u), N, _d(v), M,
                                tp, cnt, _d(y), nthreads, int(obliquity))
    return y.view(np.complex128)


def transform_cartesian(sign: int, lam: float, w, u, v, nthreads: int = 0,
                        obliquity: bool = False) -> np.ndarray:
    w = np.ascontiguousarray(w, dtype=np.complex128).view(np.float64)
    u = _f64(u).reshape(-1); v = _f64(v).reshape(-1)
    N, M = u.size // 3, v.size // 3
    y = np.zeros(2 * M, dtype=np.float64)
    lib().oracle_transform_cartesian(sign, lam, _d(w), _d(u), N, _d(v), M, _d(y), nthreads,
                                     int(obliquity))
    return y.view(np.complex128)


def normalize_polar(y):
    y = np.ascontiguousarray(y, dtype=np.complex128)
    yv = y.view(np.float64)
    amp = np.empty(y.size); phase = np.empty(y.size)
    mx = lib().oracle_normalize_polar(_d(yv), y.size, _d(amp), _d(phase))
    return amp, phase, mx


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def mc_index(seed: int, m: int, k: int, N: int) -> int:
    return int(lib().oracle_mc_index(seed, m, k, N))


def mc_shared_index(seed: int, g: int, k: int, N: int) -> int:
    return int(lib().oracle_mc_shared_index(seed, g, k, N))


def transform_mc(sign: int, lam: float, amp, phase, u, v, K: int, seed: int,
                 targets=None, nthreads: int = 0, shared: bool = False,
                 obliquity: bool = False) -> np.ndarray:
    amp = _f64(amp); phase = _f64(phase); u = _f64(u).reshape(-1); v = _f64(v).reshape(-1)
    N, M = amp.size, v.size // 3
    if targets is not None:
        t = np.ascontiguousarray(targets, dtype=np.int64)
        cnt, tp = t.size, t.ctypes.data_as(_i64p)
    else:
        cnt, tp = M, None
    y = np.zeros(2 * cnt)
    lib().oracle_transform_mc(sign, lam, _d(amp), _d(phase), _d(u), N, _d(v), M,
                              K, seed, 1 if shared else 0, tp, cnt, _d(y), nthreads,
                              int(obliquity))
    return y.view(np.complex128)


def adjoint(sign: int, lam: float, amp, theta, u, g, v, nthreads: int = 0,
            obliquity: bool = False):
    """(dL/dA, dL/dtheta) for z = sum_n A e^{i theta} G; g = dL/dRe z + i dL/dIm z."""
    amp = _f64(amp); theta = _f64(theta); u = _f64(u).reshape(-1); v = _f64(v).reshape(-1)
    g = np.ascontiguousarray(g, dtype=np.complex128).view(np.float64)
    N, M = amp.size, v.size // 3
    ga = np.empty(N); gt = np.empty(N)
    lib().oracle_adjoint(sign, lam, _d(amp), _d(theta), _d(u), N, _d(g), _d(v), M,
                         _d(ga), _d(gt), nthreads, int(obliquity))
    return ga, gt
