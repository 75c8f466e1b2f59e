"""NumPy restatement of the reference's CPU path -- TEST/BENCH INFRASTRUCTURE.

Used only by tests/ and by bench.py's cpu_baseline leg.  It restates
py/util.py (which cannot run as shipped: it imports numba, and sum_kernel
calls an undefined `f`, util.py:114; the intended alias is
f = phasor_displacement, util.py:92):

  to_polar            py/util.py:48-52   a * exp(i phi)
  from_polar          py/util.py:55-65   (|c|, angle c), angle negated if distance < 0
  norm                py/util.py:69-72   sqrt(sum(X**2, axis=1))
  phasor_displacement py/util.py:92-108  phase - direction * 2 pi delta / LAMBDA
  sum_kernel          py/util.py:111-121 np.sum over all sources
  map_sum_kernel      py/util.py:124-147 Pool(parallel).starmap over targets

Sign convention: the Python `direction=+1` subtracts 2 pi d / lambda, i.e. it
is the CUDA Direction::Backwards; CUDA sign s maps to direction = -s.  The
Python order (2 pi * delta) / lambda rounds differently from the CUDA
delta * (2 pi / lambda), so the two agree to ~1e-10 relative, not bit-wise.
"""
from __future__ import annotations

import os
from itertools import repeat
from multiprocessing import Pool

import numpy as np


def to_polar(a, phi):
    return a * np.exp(phi * 1j)


def from_polar(c, distance=1):
    amplitude = np.abs(c)
    phase = np.angle(c)
    if np.sign(distance) == -1:
        return amplitude, -phase
    return amplitude, phase


def norm(X):
    return np.sqrt(np.sum(X ** 2, axis=1))


def phasor_displacement(amplitude, phase, w, v, direction=1, lam=0.65e-6):
    delta = norm(w - v)
    next_phase = phase - direction * 2 * np.pi * delta / lam
    next_amplitude = amplitude / delta
    return to_polar(next_amplitude, next_phase)


def sum_kernel(x, w, v, direction=1, lam=0.65e-6):
    return np.sum(phasor_displacement(x[:, 0], x[:, 1], w, v, direction=direction, lam=lam))


def _sum_kernel_wrapper(v, args):
    x, w, direction, distance, lam = args
    return from_polar(sum_kernel(x, w, v, direction=direction, lam=lam), distance=distance)


def map_sum_kernel(x, y, w, v, direction=1, distance=1, lam=0.65e-6, parallel=0):
    """y[m] = polar(sum_n phasor(x_n, w_n -> v_m)); x: (N, 2), w: (N, 3), v: (M, 3)."""
    if parallel:
        workers = 4 if parallel in (True, 1) else int(parallel)
        args = repeat((x, w, direction, distance, lam))
        with Pool(workers) as p:
            results = p.starmap(_sum_kernel_wrapper, zip(v, args))
        y[:] = np.array(results)
        return
    for m in range(y.shape[0]):
        y[m, :] = from_polar(sum_kernel(x, w, v[m], direction=direction, lam=lam),
                             distance=distance)


def cpu_count() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1
