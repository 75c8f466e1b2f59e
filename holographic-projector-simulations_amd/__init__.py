"""holo_amd -- MI355X (gfx950) Huygens point-source superposition.

Python side of ``libholo.so`` (C ABI in ``include/holo.h``).  Arrays are
NumPy (host memory: the library stages them) or torch tensors on the GPU
(device pointers, ordered on torch's current stream).  There is no CPU
fallback: every call fails loudly when ``libholo.so`` is missing or no
gfx950 device is present.

Mirrors of the reference interface:
  * ``transform_full`` / ``time_transform``  -- cuda/transform.cu:147-340,
    346-406 (the C++ templates ``transform_full<Direction, Algorithm,
    shared_memory>`` and ``time_transform<Direction, ...>``).
  * ``transform`` -- the Monte Carlo estimator named in README.md:127.
  * ``Direction`` -- cuda/superposition.cu:19; ``LAMBDA`` -- cuda/macros.h:23-25.
"""
from __future__ import annotations

import ctypes
import enum
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libholo.so")

LAMBDA = 0.65e-6            # cuda/macros.h:23-25 default wavelength (m)
FLOP_PER_POINT = 29         # cuda/macros.h:71-80, the reference's pair cost
FP64 = 0
FP32_SPLIT = 1
FORWARDS = +1
BACKWARDS = -1
DEVICE_POINTERS = 0x1
MC_SHARED = 0x2
OBLIQUITY = 0x4             # amplitude x |dz|/d, matlab/holofunctions.m:226-227


class Direction(enum.IntEnum):
    """cuda/superposition.cu:19: Forwards adds 2 pi d / lambda, Backwards subtracts."""
    Forwards = +1
    Backwards = -1


class Algorithm(enum.IntEnum):
    """cuda/main.h:7 / input.h:113 (-p 1, 2, 3).  All three compute the same
    sum; here they are accepted as hints and select nothing."""
    Naive = 1
    Reduced = 2
    ReducedSharedMemory = 3


class HoloError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"holo status {status}: {message}")
        self.status = status


_lib = None
_c_double_p = ctypes.c_void_p


def lib() -> ctypes.CDLL:
    """Load libholo.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HoloError(3, f"{LIB_PATH} is missing; run build.py (or __graft_entry__.build())")
    # torch ships its own HIP runtime.  Loading it first makes libholo.so's
    # libamdhip64.so.7 dependency resolve to that same runtime, so one process
    # never holds two HIP runtimes (the second would find no GPU).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i, u32, u64, d = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                              ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double)
    L.holo_abi_version.restype = i
    L.holo_status_string.restype = ctypes.c_char_p
    L.holo_status_string.argtypes = [i]
    L.holo_last_error.restype = ctypes.c_char_p
    L.holo_create.argtypes = [ctypes.POINTER(vp), i, vp]
    L.holo_destroy.argtypes = [vp]
    L.holo_destroy.restype = None
    L.holo_set_stream.argtypes = [vp, vp]
    L.holo_synchronize.argtypes = [vp]
    L.holo_set_option.argtypes = [ctypes.c_char_p, ctypes.c_long]
    L.holo_get_option.argtypes = [ctypes.c_char_p]
    L.holo_get_option.restype = ctypes.c_long
    L.holo_transform_full.argtypes = [vp, i, i, d, vp, vp, vp, sz, vp, sz, vp, vp,
                                      ctypes.c_uint]
    L.holo_transform_full_noalloc.argtypes = [i, d, vp, vp, vp, sz, vp, sz, vp, vp,
                                              ctypes.c_uint]
    L.holo_transform_mc.argtypes = [vp, i, i, d, vp, vp, vp, sz, vp, sz, u32, u64, vp, vp,
                                    ctypes.c_uint]
    L.holo_mc_indices.argtypes = [vp, u64, u64, sz, sz, u32, vp, ctypes.c_uint]
    L.holo_normalize_polar.argtypes = [vp, vp, sz, vp, vp, vp, ctypes.c_uint]
    L.holo_plan_create.argtypes = [vp, ctypes.POINTER(vp), vp, sz, ctypes.c_uint, ctypes.c_uint]
    L.holo_plan_destroy.argtypes = [vp]
    L.holo_plan_destroy.restype = None
    L.holo_plan_info.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz),
                                 ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(d)]
    L.holo_plan_lattice.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
    L.holo_plan_transform.argtypes = [vp, vp, i, i, d, vp, vp, vp, sz, vp, vp, ctypes.c_uint]
    L.holo_plan_transform_complex.argtypes = [vp, vp, i, i, d, vp, vp, sz, vp, vp,
                                              ctypes.c_uint]
    L.holo_plan_adjoint.argtypes = [vp, vp, i, i, d, vp, vp, vp, vp, sz, vp, vp, vp,
                                    ctypes.c_uint]
    L.holo_read_dat.argtypes = [ctypes.c_char_p, vp, sz, ctypes.POINTER(sz)]
    L.holo_write_dat.argtypes = [ctypes.c_char_p, vp, sz]
    dp = ctypes.POINTER(d)
    L.holo_plan_layout.argtypes = [vp, sz, ctypes.c_uint, vp, sz, ctypes.POINTER(sz), dp,
                                   ctypes.POINTER(ctypes.c_int)]
    L.holo_plan_layout.restype = i
    L.holo_lattice_layout.argtypes = [vp, sz, ctypes.POINTER(sz), ctypes.POINTER(sz),
                                      ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_int)]
    L.holo_lattice_layout.restype = i
    L.holo_host_sincos.argtypes = [d, dp, dp]
    L.holo_host_sincos.restype = None
    L.holo_host_sqrt_inv.argtypes = [d, dp, dp]
    L.holo_host_sqrt_inv.restype = None
    for name in ("holo_create", "holo_set_stream", "holo_synchronize", "holo_set_option",
                 "holo_transform_full", "holo_transform_full_noalloc", "holo_transform_mc",
                 "holo_mc_indices",
                 "holo_normalize_polar", "holo_plan_create", "holo_plan_info", "holo_plan_lattice",
                 "holo_plan_transform", "holo_plan_transform_complex", "holo_plan_adjoint",
                 "holo_read_dat", "holo_write_dat"):
        getattr(L, name).restype = i
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        msg = lib().holo_last_error().decode(errors="replace")
        raise HoloError(status, msg)


def set_option(name: str, value: int) -> None:
    check(lib().holo_set_option(name.encode(), int(value)))


def get_option(name: str) -> int:
    return int(lib().holo_get_option(name.encode()))


# ---------------------------------------------------------------------------
# array marshalling


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _host(x, dtype=np.float64) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=dtype)
    return a


def _ptr(a) -> Optional[int]:
    if a is None:
        return None
    if _is_torch(a):
        return a.data_ptr() if a.numel() else None
    return a.ctypes.data if a.size else None


def _prep(arrays, device: bool):
    """Validate that arrays are all host or all device; return contiguous views."""
    out = []
    for a in arrays:
        if a is None:
            out.append(None)
            continue
        if device:
            import torch
            if not (_is_torch(a) and a.is_cuda):
                raise HoloError(1, "mixing device tensors and host arrays")
            if a.dtype == torch.complex128:
                a = torch.view_as_real(a)
            if a.dtype != torch.float64:
                raise HoloError(1, f"device tensors must be float64/complex128, got {a.dtype}")
            out.append(a.contiguous())
        else:
            if _is_torch(a):
                a = a.detach().cpu().numpy()
            a = np.asarray(a)
            if np.iscomplexobj(a):
                a = np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)
            out.append(np.ascontiguousarray(a, dtype=np.float64))
    return out


def _n(a, per: int) -> int:
    if a is None:
        return 0
    numel = a.numel() if _is_torch(a) else a.size
    if numel % per:
        raise HoloError(1, f"array of {numel} values is not a multiple of {per}")
    return numel // per


class Context:
    """A libholo context: one device, one stream, reusable device buffers."""

    def __init__(self, device: int = 0, stream=None):
        self._h = ctypes.c_void_p()
        s = None if stream is None else int(stream)
        check(lib().holo_create(ctypes.byref(self._h), int(device), s))
        self.device = device
        # error report of the last call made with report_error=True: the
        # relative L2 distance of an FP32-split result from the FP64
        # formulation, estimated from 256 sampled targets (include/holo.h;
        # 0 for FP64); a float, or a one-element device tensor for calls on
        # device tensors (no sync)
        self.last_rel_l2 = None

    def close(self):
        if self._h:
            lib().holo_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream) -> None:
        """Order work on `stream` (a hipStream_t handle; 0 = default stream)."""
        check(lib().holo_set_stream(self._h, int(stream) or None))

    def synchronize(self) -> None:
        check(lib().holo_synchronize(self._h))

    def _torch_stream(self):
        import torch
        self.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def _err_slot(self, report: bool, device: bool, like=None):
        """(holder, pointer) for the err_rel_l2 argument, or (None, None)."""
        if not report:
            return None, None
        if device:
            import torch
            t = torch.zeros(1, dtype=torch.float64, device=like.device)
            return t, t.data_ptr()
        c = ctypes.c_double(0.0)
        return c, ctypes.addressof(c)

    def _keep_err(self, holder, device: bool) -> None:
        if holder is not None:
            self.last_rel_l2 = holder if device else holder.value

    # -- brute force ---------------------------------------------------------
    def transform_full(self, amp, phase, u, v, sign: int = FORWARDS, lam: float = LAMBDA,
                       precision: int = FP64, obliquity: bool = False,
                       report_error: bool = False):
        """y[m] = sum_n amp_n/d exp(i(phase_n + sign 2 pi d/lam)); complex (M,).
        obliquity: each term also carries the factor |dz|/d.  report_error:
        the call also writes its error report to self.last_rel_l2."""
        extra = OBLIQUITY if obliquity else 0
        device = _is_torch(amp) and amp.is_cuda
        amp, phase, u, v = _prep([amp, phase, u, v], device)
        N, M = _n(amp, 1), _n(v, 3)
        if _n(phase, 1) != N or _n(u, 3) != N:
            raise HoloError(1, "amp, phase and u must describe the same N sources")
        if device:
            import torch
            self._torch_stream()
            y = torch.empty((M, 2), dtype=torch.float64, device=v.device)
            eh, ep = self._err_slot(report_error, True, v)
            check(lib().holo_transform_full(self._h, precision, sign, lam, _ptr(amp), _ptr(phase),
                                            _ptr(u), N, _ptr(v), M, _ptr(y), ep,
                                            DEVICE_POINTERS | extra))
            self._keep_err(eh, True)
            return torch.view_as_complex(y)
        y = np.empty(2 * M)
        eh, ep = self._err_slot(report_error, False)
        check(lib().holo_transform_full(self._h, precision, sign, lam, _ptr(amp), _ptr(phase),
                                        _ptr(u), N, _ptr(v), M, _ptr(y), ep, extra))
        self._keep_err(eh, False)
        return y.view(np.complex128)

    # -- Monte Carlo -----------------------------------------------------------
    def transform_mc(self, amp, phase, u, v, K: int, seed: int = 0, sign: int = FORWARDS,
                     lam: float = LAMBDA, precision: int = FP32_SPLIT, shared: bool = False,
                     obliquity: bool = False, report_error: bool = False):
        """Monte Carlo estimate with K draws per target (shared: per group of 1024).
        report_error: self.last_rel_l2 receives the distance from the FP64
        estimator with the same indices, estimated from 256 sampled targets."""
        extra = (MC_SHARED if shared else 0) | (OBLIQUITY if obliquity else 0)
        device = _is_torch(amp) and amp.is_cuda
        amp, phase, u, v = _prep([amp, phase, u, v], device)
        N, M = _n(amp, 1), _n(v, 3)
        if _n(phase, 1) != N or _n(u, 3) != N:
            raise HoloError(1, "amp, phase and u must describe the same N sources")
        if device:
            import torch
            self._torch_stream()
            y = torch.empty((M, 2), dtype=torch.float64, device=v.device)
            eh, ep = self._err_slot(report_error, True, v)
            check(lib().holo_transform_mc(self._h, precision, sign, lam, _ptr(amp), _ptr(phase),
                                          _ptr(u), N, _ptr(v), M, int(K), int(seed), _ptr(y), ep,
                                          DEVICE_POINTERS | extra))
            self._keep_err(eh, True)
            return torch.view_as_complex(y)
        y = np.empty(2 * M)
        eh, ep = self._err_slot(report_error, False)
        check(lib().holo_transform_mc(self._h, precision, sign, lam, _ptr(amp), _ptr(phase),
                                      _ptr(u), N, _ptr(v), M, int(K), int(seed), _ptr(y), ep,
                                      extra))
        self._keep_err(eh, False)
        return y.view(np.complex128)

    def mc_indices(self, seed: int, N: int, m0: int, count: int, K: int,
                   shared: bool = False) -> np.ndarray:
        """Indices drawn on the device: per target (or per group with shared)."""
        idx = np.empty(count * K, dtype=np.int64)
        check(lib().holo_mc_indices(self._h, int(seed), int(N), int(m0), int(count), int(K),
                                    _ptr(idx), MC_SHARED if shared else 0))
        return idx.reshape(count, K)

    # -- time_transform post-processing -------------------------------------------
    def normalize_polar(self, y):
        device = _is_torch(y) and y.is_cuda
        (yv,) = _prep([y], device)
        M = _n(yv, 2)
        if device:
            import torch
            self._torch_stream()
            amp = torch.empty(M, dtype=torch.float64, device=yv.device)
            ph = torch.empty(M, dtype=torch.float64, device=yv.device)
            mx = torch.zeros(1, dtype=torch.float64, device=yv.device)
            check(lib().holo_normalize_polar(self._h, _ptr(yv), M, _ptr(amp), _ptr(ph), _ptr(mx),
                                             DEVICE_POINTERS))
            return amp, ph, mx
        amp = np.empty(M)
        ph = np.empty(M)
        mx = ctypes.c_double(0)
        check(lib().holo_normalize_polar(self._h, _ptr(yv), M, _ptr(amp), _ptr(ph),
                                         ctypes.addressof(mx), 0))
        return amp, ph, mx.value

    def plan(self, u, tile: int = 64) -> "Plan":
        return Plan(self, u, tile)


class Plan:
    """Source geometry resident on the device (and tiled for FP32-split)."""

    def __init__(self, ctx: Context, u, tile: int = 64):
        self.ctx = ctx
        device = _is_torch(u) and u.is_cuda
        (uu,) = _prep([u], device)
        self.N = _n(uu, 3)
        self._h = ctypes.c_void_p()
        if device:
            ctx._torch_stream()
        check(lib().holo_plan_create(ctx.handle, ctypes.byref(self._h), _ptr(uu), self.N,
                                     int(tile), DEVICE_POINTERS if device else 0))

    def close(self):
        if self._h:
            lib().holo_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        n, t, tile, off = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint(), ctypes.c_double()
        check(lib().holo_plan_info(self._h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(tile),
                                   ctypes.byref(off)))
        return {"n_sources": n.value, "n_tiles": t.value, "tile": tile.value,
                "max_offset": off.value}

    def lattice(self):
        """(nx, ny) of the product grid the plan detected, or (0, 0)."""
        nx, ny = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().holo_plan_lattice(self._h, ctypes.byref(nx), ctypes.byref(ny)))
        return nx.value, ny.value

    def _out(self, like, M, device):
        if device:
            import torch
            return torch.empty((M, 2), dtype=torch.float64, device=like.device)
        return np.empty(2 * M)

    def transform(self, amp, phase, v, sign: int = FORWARDS, lam: float = LAMBDA,
                  precision: int = FP64, out=None, obliquity: bool = False,
                  report_error: bool = False):
        device = _is_torch(v) and v.is_cuda
        amp, phase, v = _prep([amp, phase, v], device)
        M = _n(v, 3)
        if _n(amp, 1) != self.N or _n(phase, 1) != self.N:
            raise HoloError(1, "weights do not match the plan's sources")
        y = out if out is not None else self._out(v, M, device)
        if device:
            self.ctx._torch_stream()
        flags = (DEVICE_POINTERS if device else 0) | (OBLIQUITY if obliquity else 0)
        eh, ep = self.ctx._err_slot(report_error, device, v)
        check(lib().holo_plan_transform(self.ctx.handle, self._h, precision, sign, lam,
                                        _ptr(amp), _ptr(phase), _ptr(v), M, _ptr(y), ep, flags))
        self.ctx._keep_err(eh, device)
        if device:
            import torch
            return torch.view_as_complex(y) if y.dim() == 2 else y
        return y.view(np.complex128)

    def transform_complex(self, w, v, sign: int = FORWARDS, lam: float = LAMBDA,
                          precision: int = FP64, obliquity: bool = False,
                          report_error: bool = False):
        device = _is_torch(v) and v.is_cuda
        w, v = _prep([w, v], device)
        M = _n(v, 3)
        if _n(w, 2) != self.N:
            raise HoloError(1, "weights do not match the plan's sources")
        y = self._out(v, M, device)
        if device:
            self.ctx._torch_stream()
        flags = (DEVICE_POINTERS if device else 0) | (OBLIQUITY if obliquity else 0)
        eh, ep = self.ctx._err_slot(report_error, device, v)
        check(lib().holo_plan_transform_complex(self.ctx.handle, self._h, precision, sign, lam,
                                                _ptr(w), _ptr(v), M, _ptr(y), ep, flags))
        self.ctx._keep_err(eh, device)
        if device:
            import torch
            return torch.view_as_complex(y)
        return y.view(np.complex128)

    def adjoint(self, g, amp, theta, u, sign: int = FORWARDS, lam: float = LAMBDA,
                precision: int = FP64, obliquity: bool = False, report_error: bool = False):
        """(dL/damp, dL/dtheta) for z = transform(amp, theta; u -> plan points).
        report_error: ctx.last_rel_l2 receives the report of the field
        b = sum_m g_m conj(G_mn) at the points u."""
        device = _is_torch(u) and u.is_cuda
        g, amp, theta, u = _prep([g, amp, theta, u], device)
        N = _n(u, 3)
        if _n(g, 2) != self.N or _n(amp, 1) != N or _n(theta, 1) != N:
            raise HoloError(1, "shapes of g / amp / theta / u do not match")
        if device:
            import torch
            self.ctx._torch_stream()
            ga = torch.empty(N, dtype=torch.float64, device=u.device)
            gt = torch.empty(N, dtype=torch.float64, device=u.device)
        else:
            ga, gt = np.empty(N), np.empty(N)
        eh, ep = self.ctx._err_slot(report_error, device, u)
        check(lib().holo_plan_adjoint(self.ctx.handle, self._h, precision, sign, lam, _ptr(g),
                                      _ptr(amp), _ptr(theta), _ptr(u), N, _ptr(ga), _ptr(gt), ep,
                                      (DEVICE_POINTERS if device else 0)
                                      | (OBLIQUITY if obliquity else 0)))
        self.ctx._keep_err(eh, device)
        return ga, gt


def plan_layout(u, tile: int = 64):
    """Host-side tiling of positions u (no device): (perm, n_tiles, max_offset, planar)."""
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    N = u.size // 3
    nt, off, pl = ctypes.c_size_t(), ctypes.c_double(), ctypes.c_int()
    check(lib().holo_plan_layout(_ptr(u), N, int(tile), None, 0, ctypes.byref(nt),
                                 ctypes.byref(off), ctypes.byref(pl)))
    perm = np.empty(max(nt.value * (tile or 64), 1), dtype=np.int32)
    check(lib().holo_plan_layout(_ptr(u), N, int(tile), _ptr(perm), perm.size, ctypes.byref(nt),
                                 ctypes.byref(off), ctypes.byref(pl)))
    return perm[:nt.value * (tile or 64)], nt.value, off.value, bool(pl.value)


def lattice_layout(u):
    """Host-side lattice detection of a plan over positions u (no device):
    (nx, ny, tail, tiles16) -- the lattice (0, 0 if none), the sources after
    it summed exactly, and whether 16-row tiles are built."""
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    nx, ny, tail, t16 = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
    check(lib().holo_lattice_layout(_ptr(u), u.size // 3, ctypes.byref(nx), ctypes.byref(ny),
                                    ctypes.byref(tail), ctypes.byref(t16)))
    return nx.value, ny.value, tail.value, bool(t16.value)


def host_sincos(x: float):
    """Host build of the FP64 kernel's sin/cos: (sin x, cos x)."""
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().holo_host_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


# ---------------------------------------------------------------------------
# .dat files (cuda/util.h:248-337): headerless little-endian float64


def read_dat(path: str) -> np.ndarray:
    n = ctypes.c_size_t()
    check(lib().holo_read_dat(path.encode(), None, 0, ctypes.byref(n)))
    buf = np.empty(n.value)
    check(lib().holo_read_dat(path.encode(), _ptr(buf), n.value, ctypes.byref(n)))
    return buf


def write_dat(path: str, values) -> None:
    a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    check(lib().holo_write_dat(path.encode(), _ptr(a), a.size))


def parse_json(lines) -> dict:
    """The out.json lines of a run, grouped by the first letter of "phasor"
    (x, y, z; u, v, w stay empty), as py/util.py:686-698."""
    import json
    params = {k: [] for k in "xyzuvw"}
    for line in lines:
        if isinstance(line, bytes):
            line = line.decode()
        if line.strip():
            p = json.loads(line)
            k = p["phasor"][:1]
            if k not in "xyz":
                raise ValueError(f"unexpected phasor key {p['phasor']!r}")
            params[k].append(p)
    return params


def parse_file(dir: str = "../tmp", zipfilename: Optional[str] = "out.zip", prefix: str = "out",
               read_data: bool = True) -> Tuple[dict, ...]:
    """A run's metadata and arrays, as py/util.py:700-744: from the archive
    dir/zipfilename when it exists (the reference's layout), else from the
    files in dir.  params[k] lists the out.json records of phasor k;
    data[k] the (len, 2) [amp, phase] arrays of each record and data[p] its
    (len, 3) positions (p = first letter of "pos").  A record whose files are
    shorter than its "len" gets len 0 and no arrays, as in the reference.
    Returns (params,) when read_data is False."""
    import zipfile
    archive = None
    path = os.path.join(dir, zipfilename) if zipfilename else None
    if path and os.path.exists(path):
        archive = zipfile.ZipFile(path)

    def read_bytes(name: str, count: int, precision: int) -> np.ndarray:
        if precision not in (4, 8, 16):
            raise NotImplementedError(f"precision {precision}")
        if archive is not None:
            size = archive.getinfo(name).file_size
        else:
            size = os.path.getsize(os.path.join(dir, name))
        if size < count * precision:
            raise EOFError(f"{name}: {size} bytes < {count} x {precision}")
        if archive is not None:
            with archive.open(name) as f:
                raw = f.read(count * precision)
        else:
            with open(os.path.join(dir, name), "rb") as f:
                raw = f.read(count * precision)
        return np.frombuffer(raw, dtype=f"<f{precision}", count=count).copy()

    try:
        if archive is not None:
            with archive.open(prefix + ".json") as f:
                params = parse_json(f.readlines())
        else:
            with open(os.path.join(dir, prefix + ".json")) as f:
                params = parse_json(f.readlines())
        if not read_data:
            return (params,)
        data = {k: [] for k in "xyzuvw"}
        for k in "xyz":
            for p in params[k]:
                k2 = p["pos"][:1]
                p["len"] = int(p["len"])
                try:
                    amp = read_bytes(p["phasor"] + "_amp.dat", p["len"], p["precision"])
                    phase = read_bytes(p["phasor"] + "_phase.dat", p["len"], p["precision"])
                    pos = read_bytes(p["pos"] + ".dat", p["len"] * p["dims"], p["precision"])
                except (EOFError, KeyError, FileNotFoundError):
                    p["len"] = 0
                    continue
                data[k].append(np.stack([amp, phase], axis=1))
                data[k2].append(pos.reshape(-1, p["dims"]))
        return params, data
    finally:
        if archive is not None:
            archive.close()


# ---------------------------------------------------------------------------
# Mirrors of the reference's C++ templates (cuda/transform.cu)

_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def transform_full(x, u, v, direction: Direction = Direction.Forwards,
                   algorithm: Algorithm = Algorithm.Reduced, shared_memory: bool = False,
                   lam: float = LAMBDA, precision: int = FP64,
                   ctx: Optional[Context] = None) -> np.ndarray:
    """transform_full<direction, algorithm, shared_memory>(x, u, v, p)
    (cuda/transform.cu:147-151).  x: (N, 2) polar [amp, phase]; u: 3N; v: 3M.
    Returns the unnormalised complex field at the M targets.  The reference
    also returns padded batch entries; batching here is internal, so none
    are appended (the C++ mirror in include/holo/transform.hpp keeps them).
    The reference asserts u[2] != v[2] (cuda/transform.cu:193)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    u = np.asarray(u, dtype=np.float64).reshape(-1)
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    if u.size and v.size and u[2] == v[2]:
        raise HoloError(1, "first source and first target share a z offset "
                           "(cuda/transform.cu:193 asserts u[2] != v[2])")
    c = ctx or default_context()
    return c.transform_full(x[:, 0], x[:, 1], u, v, int(direction), lam, precision)


def time_transform(x, u, v, direction: Direction = Direction.Forwards,
                   algorithm: Algorithm = Algorithm.Reduced, lam: float = LAMBDA,
                   precision: int = FP64, ctx: Optional[Context] = None
                   ) -> Tuple[np.ndarray, float]:
    """time_transform<direction>(...) (cuda/transform.cu:346-406) without the
    reference wave: returns ((M, 2) polar [amp, phase] normalised to max 1,
    runtime in seconds)."""
    import time
    t0 = time.perf_counter()
    y = transform_full(x, u, v, direction, algorithm, lam=lam, precision=precision, ctx=ctx)
    amp, phase, _ = (ctx or default_context()).normalize_polar(y)
    dt = time.perf_counter() - t0
    return np.stack([amp, phase], axis=1), dt


def transform(x, u, v, K: int, seed: int = 0, direction: Direction = Direction.Forwards,
              lam: float = LAMBDA, precision: int = FP32_SPLIT, batched: bool = False,
              ctx: Optional[Context] = None) -> np.ndarray:
    """Monte Carlo transform(): K uniform source draws per target ("True MC"),
    or per group of 1024 targets with batched=True ("Batched MC"), scaled N/K."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    c = ctx or default_context()
    return c.transform_mc(x[:, 0], x[:, 1], u, v, K, seed, int(direction), lam, precision,
                          shared=batched)
