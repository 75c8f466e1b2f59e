"""The FP64 kernels' own sin/cos (Cody-Waite reduction modulo pi/1024, a
correctly rounded table of (sin, cos)(j pi/1024) and short Taylor polynomials,
holo_math.hpp) through its host build: accuracy against mpmath and glibc over
the superposition's argument range and the reduction's edges, and exact
odd/even symmetry.  tests/test_gpu_parity.py::test_fp64_device_sincos_bits
checks that the device produces the same bits."""
import math

import mpmath as mp
import numpy as np


def test_sincos_accuracy_random(pkg):
    g = np.random.default_rng(0)
    xs = np.concatenate([g.uniform(-4, 4, 20000), g.uniform(0, 2e7, 20000),
                         -g.uniform(0, 2e7, 20000), g.uniform(0, 1e9, 5000)])
    worst = 0.0
    for x in xs:
        s, c = pkg.host_sincos(x)
        worst = max(worst, abs(s - math.sin(x)), abs(c - math.cos(x)))
    assert worst <= 2.5e-16  # measured: 2.2e-16 (one ulp of values near 1)


def test_sincos_reduction_edges_against_mpmath(pkg):
    """Arguments next to multiples of pi/2 (where the reduced value or the
    result is tiny), of pi (sign flips) and of pi/1024 (table steps), checked
    at 40 digits."""
    mp.mp.dps = 40
    for k in (1, 2, 3, 7, 1000, 123457, 2 ** 21 + 3):
        for base in (k * math.pi / 2, k * math.pi, k * math.pi / 1024, (k + 0.5) * math.pi / 1024):
            for ulps in (-3, -1, 0, 1, 3):
                x = base + ulps * math.ulp(base)
                s, c = pkg.host_sincos(x)
                assert abs(s - float(mp.sin(mp.mpf(x)))) <= 2.5e-16
                assert abs(c - float(mp.cos(mp.mpf(x)))) <= 2.5e-16


def test_sincos_odd_even_bitwise(pkg):
    """sin(-x) = -sin(x) and cos(-x) = cos(x) bit for bit: the FP64 field of
    Backwards at phase -phi is then exactly the conjugate of Forwards."""
    g = np.random.default_rng(2)
    for x in np.concatenate([g.uniform(0, 8, 3000), g.uniform(0, 3e7, 3000)]):
        s, c = pkg.host_sincos(x)
        sm, cm = pkg.host_sincos(-x)
        assert sm == -s and cm == c


def test_sincos_known_values(pkg):
    assert pkg.host_sincos(0.0) == (0.0, 1.0)
    s, c = pkg.host_sincos(-2 * math.pi)  # the reference's known answer argument
    assert s == 2.4492935982947064e-16 and c == 1.0


def test_sqrt_inv_host(pkg):
    import ctypes
    g = np.random.default_rng(1)
    d, inv = ctypes.c_double(), ctypes.c_double()
    for s in 10.0 ** g.uniform(-8, 1, 20000):
        pkg.lib().holo_host_sqrt_inv(s, ctypes.byref(d), ctypes.byref(inv))
        assert d.value == math.sqrt(s)
        assert abs(inv.value * d.value - 1) <= 4e-16
