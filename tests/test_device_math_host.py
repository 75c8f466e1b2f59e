"""The FP64 kernels' own sin/cos (Cody-Waite reduction modulo pi/1024, a
correctly rounded table of (sin, cos)(j pi/1024), sin r = r - r^3/6 and the
minimax cos r = 1 + C2 r^2 on |r| <= pi/2048, holo_math.hpp) through its host
build: accuracy against mpmath and glibc over the superposition's argument
range and the reduction's edges, and exact odd/even symmetry.  The bound is
the cos term's 3.96e-14 plus rounding; the Taylor coefficient -1/2 in its place
(2.3e-13) or a wrong table step fails it, and the size of the minimax error is
pinned from both sides next to a table step.
tests/test_gpu_parity.py::test_fp64_device_sincos_bits checks that the device
produces the same bits."""
import math

import mpmath as mp
import numpy as np

TOL = 4.2e-14  # 3.96e-14 (the cos term) + rounding


def test_sincos_accuracy_random(pkg):
    g = np.random.default_rng(0)
    xs = np.concatenate([g.uniform(-4, 4, 20000), g.uniform(0, 2e7, 20000),
                         -g.uniform(0, 2e7, 20000), g.uniform(0, 1e9, 5000)])
    worst = 0.0
    for x in xs:
        s, c = pkg.host_sincos(x)
        worst = max(worst, abs(s - math.sin(x)), abs(c - math.cos(x)))
    assert worst <= TOL


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
                assert abs(s - float(mp.sin(mp.mpf(x)))) <= TOL
                assert abs(c - float(mp.cos(mp.mpf(x)))) <= TOL


def test_sincos_minimax_error_size(pkg):
    """At r = pi/2048 past a table step the cos term's error is close to its
    bound (3.96e-14; the Taylor -1/2 gives 2.3e-13 there, the r^4 term ~1e-16),
    and it vanishes at the step itself."""
    mp.mp.dps = 40
    for j in (1, 77, 512, 1500):
        x = float((j + mp.mpf(0.5)) * mp.pi / 1024)  # r = +-pi/2048
        s, c = pkg.host_sincos(x)
        e = math.hypot(s - float(mp.sin(mp.mpf(x))), c - float(mp.cos(mp.mpf(x))))
        assert 3.0e-14 <= e <= TOL
        x = float(j * mp.pi / 1024)  # r ~ 0
        s, c = pkg.host_sincos(x)
        assert abs(s - float(mp.sin(mp.mpf(x)))) <= 3e-16
        assert abs(c - float(mp.cos(mp.mpf(x)))) <= 3e-16


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
