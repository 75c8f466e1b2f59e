"""Synthetic inputs of the benchmark workloads (planes of points, seeded fields).

Geometry follows the reference's centred lattice (cuda/init.h:109-114, 139-140):
point (i, j) of an n_x x n_y plane sits at ((i - (n_x-1)/2) p, (j - (n_y-1)/2) p, z)
and has flat index i*n_y + j (x slow, cuda/macros.h:59 Ix2D).  Positions are
AoS xyz float64 (cuda/macros.h:48-56).

Random fields use ``numpy.random.Generator(numpy.random.Philox(seed))``;
amplitudes are drawn first, then phases.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

LAMBDA_HENE = 632.8e-9  # wavelength of every benchmark workload (m)


def grid(nx: int, ny: int, pitch: float, z: float) -> np.ndarray:
    """(nx*ny, 3) centred lattice at height z."""
    ix = (np.arange(nx) - (nx - 1) / 2.0) * pitch
    iy = (np.arange(ny) - (ny - 1) / 2.0) * pitch
    X, Y = np.meshgrid(ix, iy, indexing="ij")
    u = np.empty((nx * ny, 3))
    u[:, 0] = X.reshape(-1)
    u[:, 1] = Y.reshape(-1)
    u[:, 2] = z
    return u


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def gaussian_spots(n: int, count: int = 64, sigma: float = 3.0, seed: int = 0) -> np.ndarray:
    """n x n image of `count` Gaussian spots (peak 1, std `sigma` px) on zero,
    composed by pointwise max so the peak stays 1.  Centres are uniform in
    [3 sigma, n - 3 sigma).  Returned flattened, index i*n + j."""
    g = rng(seed)
    cx = g.uniform(3 * sigma, n - 3 * sigma, count)
    cy = g.uniform(3 * sigma, n - 3 * sigma, count)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    img = np.zeros((n, n))
    for a, b in zip(cx, cy):
        np.maximum(img, np.exp(-((i - a) ** 2 + (j - b) ** 2) / (2 * sigma ** 2)), out=img)
    return img.reshape(-1)


@dataclass
class Stage:
    """One transform: sources (amp, phase at u) -> targets v with a sign."""
    amp: Optional[np.ndarray]
    phase: Optional[np.ndarray]
    u: np.ndarray
    v: np.ndarray
    sign: int
    lam: float = LAMBDA_HENE

    @property
    def pairs(self) -> int:
        return (self.u.size // 3) * (self.v.size // 3)


@dataclass
class Workload:
    name: str
    stages: List[Stage]
    extra: dict = field(default_factory=dict)

    @property
    def pairs(self) -> int:
        return sum(s.pairs for s in self.stages)


def plain_small(seed: int = 0) -> Workload:
    """32x32 sources (z=0, 10 um, amp U[0,1), phase U[0,2pi)) -> 64x64 targets
    (z=0.35 m, 7.56 um), sign +1: the CPU-runnable brute-force case."""
    g = rng(seed)
    amp = g.random(32 * 32)
    phase = g.random(32 * 32) * 2 * np.pi
    return Workload("plain_32x32_to_64x64", [Stage(amp, phase, grid(32, 32, 10e-6, 0.0),
                                                   grid(64, 64, 7.56e-6, 0.35), +1)])


def chain_xyz(seed: int = 0, n_x: int = 512, n_y: int = 1024, n_z: int = 512) -> Workload:
    """x (512^2, z=0, 10 um, amp = 64 Gaussian spots, phase U[0,2pi)) -> y
    (1024^2, z=0.35 m, 7.56 um, sign +1) -> z (512^2, z=0, 10 um, sign -1).
    The y stage's amp/phase are the normalised polar field of stage 1."""
    g = rng(seed)
    amp = gaussian_spots(n_x, 64, 3.0, seed)
    phase = g.random(n_x * n_x) * 2 * np.pi
    u = grid(n_x, n_x, 10e-6, 0.0)
    v = grid(n_y, n_y, 7.56e-6, 0.35)
    w = grid(n_z, n_z, 10e-6, 0.0)
    return Workload(f"chain_x{n_x}^2_y{n_y}^2_z{n_z}^2",
                    [Stage(amp, phase, u, v, +1), Stage(None, None, v, w, -1)])


def reference_plane(n: int, width: float, z: float, aspect: float = 1.0,
                    offset=(0.0, 0.0)) -> np.ndarray:
    """(n, 3) positions from the reference's plane generator without jitter
    (cuda/init.h:96-157): x = floor(sqrt(n aspect)) rows of y = n // x points,
    x slow, centred on offset, pitches width / x and (width / aspect) / y; the
    n - x y points left over are copies of point 0."""
    x = int(np.sqrt(n * aspect))
    y = n // x
    height = width / aspect
    dx, dy = width / x, height / y
    u = np.empty((n, 3))
    i, j = np.meshgrid(np.arange(x), np.arange(y), indexing="ij")
    u[:x * y, 0] = (i * dx - (width - dx) / 2 + offset[0]).ravel()
    u[:x * y, 1] = (j * dy - (height - dy) / 2 + offset[1]).ravel()
    u[:x * y, 2] = z
    u[x * y:] = u[0]
    return u


def plane_points(seed: int = 0, n_src: int = 1024 * 1080, n_tgt: int = 1024) -> Workload:
    """n_src sources from the reference's plane generator, 10.24 mm wide at z=0
    (amp U[0,1), phase U[0,2pi)) -> n_tgt^2 targets (z=0.35 m, 7.56 um), sign +1;
    n_src = 1024 * 1080 is the object count of the reference's MATLAB
    comparison (py/experiment.py:179)."""
    g = rng(seed)
    amp = g.random(n_src)
    phase = g.random(n_src) * 2 * np.pi
    return Workload(f"plane_{n_src}_points_to_{n_tgt}^2",
                    [Stage(amp, phase, reference_plane(n_src, 1024 * 10e-6, 0.0),
                           grid(n_tgt, n_tgt, 7.56e-6, 0.35), +1)])


def plane_1024(seed: int = 0, n_src: int = 1024, n_tgt: int = 1024) -> Workload:
    """1024^2 sources (z=0, 10 um, amp U[0,1), phase U[0,2pi)) -> 1024^2 targets
    (z=0.35 m, 7.56 um), sign +1: the FP64 vs FP32-split and Monte Carlo case."""
    g = rng(seed)
    amp = g.random(n_src * n_src)
    phase = g.random(n_src * n_src) * 2 * np.pi
    return Workload(f"plane_{n_src}^2_to_{n_tgt}^2",
                    [Stage(amp, phase, grid(n_src, n_src, 10e-6, 0.0),
                           grid(n_tgt, n_tgt, 7.56e-6, 0.35), +1)])


def hologram(seed: int = 0, n_proj: int = 1024, n_img: int = 512, z: float = 0.35) -> Workload:
    """Phase-only hologram: theta on a 1024^2 projector (z=0, 7.56 um, amp 1,
    theta U[0,2pi)) -> 512^2 image plane (z=0.35 m, 10 um); target intensity
    = 64 Gaussian spots (sigma 3 px)."""
    g = rng(seed)
    theta = g.random(n_proj * n_proj) * 2 * np.pi
    target = gaussian_spots(n_img, 64, 3.0, seed + 1)
    st = Stage(np.ones(n_proj * n_proj), theta, grid(n_proj, n_proj, 7.56e-6, 0.0),
               grid(n_img, n_img, 10e-6, z), +1)
    return Workload(f"hologram_{n_proj}^2_to_{n_img}^2", [st], {"target": target})


def stride_subset(M: int, count: int = 4096) -> np.ndarray:
    """Fixed target subset m = k * (M / count), k = 0..count-1."""
    count = min(count, M)
    return (np.arange(count, dtype=np.int64) * M) // count
