"""CPU tests that pin the oracle (oracle/holo_oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement -- a
dropped term, a wrong sign, a transposed operand, a different rounding
order -- fails at least one of them:

* the reference's own known answer (py/test_util.py:63-64);
* an exact-rational pure-Python restatement of d and arg, bit for bit;
* mpmath (50 digits) closed forms and sums on tiny problems;
* exact symmetries (conjugation, reciprocity) and linearity;
* Random123 known-answer vectors for Philox4x32-10;
* finite differences and a dense complex matrix for the adjoint.
"""
import math
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest

LAM_REF = 0.65e-6      # cuda/macros.h:23-25, py/util.py LAMBDA
LAM = 632.8e-9


def rand_geom(seed, N, M, z=0.35, spread=5e-3):
    g = np.random.default_rng(seed)
    u = np.zeros((N, 3))
    u[:, :2] = (g.random((N, 2)) - 0.5) * spread
    v = np.zeros((M, 3))
    v[:, :2] = (g.random((M, 2)) - 0.5) * spread
    v[:, 2] = z
    amp = g.random(N)
    phase = g.random(N) * 2 * np.pi
    return amp, phase, u, v


# --------------------------------------------------------------------------
# the reference's known answer


@pytest.mark.parametrize("sign, phi_expected", [(-1, 2.4492935982947064e-16),
                                                (+1, -2.4492935982947064e-16)])
def test_reference_known_answer(oracle, sign, phi_expected):
    """py/test_util.py:49-64: source at 0 (amp 1, phase 0), target at (0,0,lambda),
    Python direction=+1 (= CUDA Backwards) -> a == 1/lambda, phi == 2.449e-16."""
    re, im, d, arg = oracle.pair(sign, LAM_REF, 1.0, 0.0, [0, 0, 0], [0, 0, LAM_REF])
    assert d == LAM_REF
    assert math.hypot(re, im) == 1538461.5384615385
    assert math.atan2(im, re) == phi_expected


def test_wavenumber_is_reference_constant(oracle):
    """TWO_PI_OVER_LAMBDA = (2 * M_PI) / LAMBDA (cuda/macros.h:29-31)."""
    for lam in (LAM_REF, LAM, 1e-6, 5.32e-7):
        assert oracle.wavenumber(lam) == (2 * math.pi) / lam


# --------------------------------------------------------------------------
# exact-rational restatement of the pinned operation order


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))  # correctly rounded


def exact_d_arg(sign, K, phi, u, v):
    dx, dy, dz = v[0] - u[0], v[1] - u[1], v[2] - u[2]
    s = fma(dz, dz, fma(dy, dy, dx * dx))
    d = math.sqrt(s)
    return d, fma(d if sign > 0 else -d, K, phi)


def unfused_d_arg(sign, K, phi, u, v):
    dx, dy, dz = v[0] - u[0], v[1] - u[1], v[2] - u[2]
    d = math.sqrt(dx * dx + dy * dy + dz * dz)
    return d, phi + (d if sign > 0 else -d) * K


@pytest.mark.parametrize("sign", [+1, -1])
def test_pair_bitwise_against_exact_rational(oracle, sign):
    amp, phase, u, v = rand_geom(1, 400, 5)
    K = oracle.wavenumber(LAM)
    unfused_mismatch = 0
    for n in range(u.shape[0]):
        for m in range(v.shape[0]):
            re, im, d, arg = oracle.pair(sign, LAM, amp[n], phase[n], u[n], v[m])
            d0, a0 = exact_d_arg(sign, K, phase[n], u[n], v[m])
            assert d == d0 and arg == a0, (n, m)
            # amplitude and phasor: a/d, r cos(arg), r sin(arg); gcc merges
            # sin and cos into glibc's sincos, which may differ by an ulp
            r = amp[n] / d
            assert abs(re - r * math.cos(arg)) <= 2 * math.ulp(r)
            assert abs(im - r * math.sin(arg)) <= 2 * math.ulp(r)
            d1, a1 = unfused_d_arg(sign, K, phase[n], u[n], v[m])
            unfused_mismatch += (d1 != d) or (a1 != arg)
    # the test distinguishes the pinned order from the naive one
    assert unfused_mismatch > 0


# --------------------------------------------------------------------------
# mpmath closed forms


def test_single_source_closed_form(oracle):
    mp.mp.dps = 50
    amp, phase, u, v = rand_geom(2, 6, 6)
    for sign in (+1, -1):
        for n in range(6):
            y = oracle.transform_full(sign, LAM, amp[n:n + 1], phase[n:n + 1], u[n:n + 1], v)
            for m in range(6):
                dx = [mp.mpf(v[m, i]) - mp.mpf(u[n, i]) for i in range(3)]
                d = mp.sqrt(sum(x * x for x in dx))
                ref = mp.mpf(amp[n]) / d * mp.expj(mp.mpf(phase[n]) + sign * 2 * mp.pi * d / mp.mpf(LAM))
                # |y| = a/d to rounding; the phase carries FP64's intrinsic
                # ~5e-10 rad error of a 3.5e6 rad argument
                assert abs(abs(y[m]) - float(abs(ref))) <= 4e-16 * abs(y[m])
                assert abs(y[m] - complex(ref)) <= 2e-9 * abs(y[m])


def test_sum_against_mpmath(oracle):
    mp.mp.dps = 50
    amp, phase, u, v = rand_geom(3, 16, 8)
    for sign in (+1, -1):
        y = oracle.transform_full(sign, LAM, amp, phase, u, v)
        for m in range(v.shape[0]):
            tot = mp.mpc(0)
            for n in range(u.shape[0]):
                dx = [mp.mpf(v[m, i]) - mp.mpf(u[n, i]) for i in range(3)]
                d = mp.sqrt(sum(x * x for x in dx))
                tot += mp.mpf(amp[n]) / d * mp.expj(mp.mpf(phase[n]) + sign * 2 * mp.pi * d / mp.mpf(LAM))
            scale = float(sum(mp.mpf(a) for a in amp)) / 0.35
            assert abs(y[m] - complex(tot)) <= 4e-9 * scale


def test_mpmath_rejects_plausible_mistakes(oracle):
    """A flipped sign or dropped 1/d term moves the sum by O(1)."""
    amp, phase, u, v = rand_geom(4, 16, 4)
    yf = oracle.transform_full(+1, LAM, amp, phase, u, v)
    yb = oracle.transform_full(-1, LAM, amp, phase, u, v)
    assert np.linalg.norm(yf - yb) > 0.1 * np.linalg.norm(yf)
    d = np.linalg.norm(v[0] - u, axis=1)
    no_div = np.sum(amp * np.exp(1j * (phase + 2 * np.pi * d / LAM)))
    assert abs(no_div - yf[0]) > 0.5 * abs(yf[0])


# --------------------------------------------------------------------------
# exact symmetries and linearity


def test_conjugation_symmetry_exact(oracle):
    """y_Backwards(phi) == conj(y_Forwards(-phi)) bit for bit (fma is odd)."""
    amp, phase, u, v = rand_geom(5, 64, 32)
    yb = oracle.transform_full(-1, LAM, amp, phase, u, v)
    yf = oracle.transform_full(+1, LAM, amp, -phase, u, v)
    assert np.array_equal(yb, np.conj(yf))


def test_reciprocity_exact(oracle):
    """d(u, v) == d(v, u): a single pair is symmetric in source and target."""
    amp, phase, u, v = rand_geom(6, 20, 20)
    for n in range(20):
        _, _, d1, a1 = oracle.pair(+1, LAM, 1.0, 0.0, u[n], v[n])
        _, _, d2, a2 = oracle.pair(+1, LAM, 1.0, 0.0, v[n], u[n])
        assert d1 == d2 and a1 == a2


def test_linearity(oracle):
    amp, phase, u, v = rand_geom(7, 50, 20)
    y1 = oracle.transform_full(+1, LAM, amp, phase, u, v)
    y2 = oracle.transform_full(+1, LAM, 2.5 * amp, phase, u, v)
    assert np.allclose(y2, 2.5 * y1, rtol=1e-14, atol=0)
    # superposition of disjoint source sets
    ya = oracle.transform_full(+1, LAM, amp[:20], phase[:20], u[:20], v)
    yb = oracle.transform_full(+1, LAM, amp[20:], phase[20:], u[20:], v)
    assert np.allclose(ya + yb, y1, rtol=1e-13, atol=1e-13 * np.abs(y1).max())


def test_unit_phasors_sum_to_n(oracle):
    """TEST_CONST_PHASE invariant (cuda/transform.cu:276-291) restated: with
    a_n = d_n and phi_n = -sign d_n K every term is ~1 and the sum is ~N."""
    amp, phase, u, v = rand_geom(8, 500, 1)
    d = np.linalg.norm(v[0] - u, axis=1)
    K = oracle.wavenumber(LAM)
    for sign in (+1, -1):
        y = oracle.transform_full(sign, LAM, d, -sign * d * K, u, v)
        assert abs(y[0] - 500) < 1e-6


def test_target_subset_matches_full(oracle):
    amp, phase, u, v = rand_geom(9, 40, 30)
    y = oracle.transform_full(+1, LAM, amp, phase, u, v)
    sub = np.array([29, 0, 7, 7])
    assert np.array_equal(oracle.transform_full(+1, LAM, amp, phase, u, v, targets=sub), y[sub])


def test_normalize_polar(oracle):
    amp, phase, u, v = rand_geom(10, 40, 30)
    y = oracle.transform_full(+1, LAM, amp, phase, u, v)
    a, p, mx = oracle.normalize_polar(y)
    assert mx == pytest.approx(np.abs(y).max(), rel=4e-16)  # hypot may differ by an ulp
    assert a.max() == pytest.approx(1.0, abs=1e-15)
    assert np.all(p > -np.pi - 1e-15) and np.all(p <= np.pi)
    assert np.allclose(a * np.exp(1j * p), y / mx, rtol=1e-14, atol=1e-15)
    # below 1e-6 nothing is divided (cuda/transform.cu:114-117)
    small = y * 1e-12 / mx
    a2, _, mx2 = oracle.normalize_polar(small)
    assert mx2 < 1e-6 and np.allclose(a2, np.abs(small), rtol=1e-14)


# --------------------------------------------------------------------------
# Philox4x32-10 and Monte Carlo indices

RANDOM123_KAT = [  # Random123 kat_vectors, philox4x32 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr, key, out", RANDOM123_KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert oracle.philox4x32_10(ctr, key) == out


def py_philox(ctr, key):
    c = list(ctr)
    k = list(key)
    for r in range(10):
        if r:
            k[0] = (k[0] + 0x9E3779B9) & 0xFFFFFFFF
            k[1] = (k[1] + 0xBB67AE85) & 0xFFFFFFFF
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k[0]) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k[1]) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
    return c


def py_mc_index(seed, m, k, N):
    o = py_philox([k >> 1, m & 0xFFFFFFFF, m >> 32, 0x4d43],
                  [seed & 0xFFFFFFFF, seed >> 32])
    r = (o[3] << 32 | o[2]) if (k & 1) else (o[1] << 32 | o[0])
    return (r * N) >> 64


def test_mc_index_matches_python_restatement(oracle):
    g = np.random.default_rng(11)
    for _ in range(300):
        seed = int(g.integers(0, 2 ** 63))
        m = int(g.integers(0, 2 ** 31))
        k = int(g.integers(0, 2 ** 20))
        N = int(g.integers(1, 2 ** 31))
        assert oracle.mc_index(seed, m, k, N) == py_mc_index(seed, m, k, N)


def test_lemire_two_product_form_is_exact():
    """The kernels map r = hi 2^32 + lo into [0, N) as
    (hi N + (lo N >> 32)) >> 32 with 32 x 32 -> 64-bit products
    (holo_device.hpp, lemire_index); for N < 2^32 that equals the 128-bit
    (r N) >> 64 the oracle forms.  Extreme words (all ones, carries out of the
    low product) and random ones; a form that drops lo N, or adds it without
    the shift, fails."""
    g = np.random.default_rng(21)
    words = [0, 1, 2 ** 32 - 1, 2 ** 31, 2 ** 64 - 1, 2 ** 64 - 2 ** 32, 2 ** 32,
             0xFFFFFFFF00000001] + [int(x) for x in g.integers(0, 2 ** 63, 3000)] + \
            [int(x) | (1 << 63) for x in g.integers(0, 2 ** 63, 3000)]
    Ns = [1, 2, 7, 1000, 2 ** 20, 2 ** 31 - 1, 2 ** 32 - 1, 3 * 2 ** 29 + 12345] + \
         [int(x) for x in g.integers(1, 2 ** 32, 200)]
    dropped = 0
    for i, r in enumerate(words):
        hi, lo = r >> 32, r & 0xFFFFFFFF
        for N in Ns[i % 7::7] + Ns[:8]:
            t = hi * N + ((lo * N) >> 32)
            assert t < 2 ** 64
            assert t >> 32 == (r * N) >> 64
            dropped += (hi * N) >> 32 != (r * N) >> 64
    assert dropped > 0  # the low product matters on some of these inputs


def test_mc_shared_index_matches_python_restatement(oracle):
    """Batched MC: group g's draws use counter (k/2, g, 0x80000000, 0x4d43)."""
    g = np.random.default_rng(12)
    for _ in range(200):
        seed = int(g.integers(0, 2 ** 63))
        grp = int(g.integers(0, 2 ** 21))
        k = int(g.integers(0, 2 ** 20))
        N = int(g.integers(1, 2 ** 31))
        o = py_philox([k >> 1, grp, 0x80000000, 0x4d43], [seed & 0xFFFFFFFF, seed >> 32])
        r = (o[3] << 32 | o[2]) if (k & 1) else (o[1] << 32 | o[0])
        assert oracle.mc_shared_index(seed, grp, k, N) == (r * N) >> 64
    # shared and per-target streams differ
    assert [oracle.mc_shared_index(5, 0, k, 1 << 30) for k in range(8)] != \
           [oracle.mc_index(5, 0, k, 1 << 30) for k in range(8)]


def test_mc_shared_estimator_unbiased(oracle):
    amp, phase, u, v = rand_geom(18, 64, 3)
    y = oracle.transform_full(+1, LAM, amp, phase, u, v)
    K, S = 16, 400
    est = np.stack([oracle.transform_mc(+1, LAM, amp, phase, u, v, K, seed, shared=True)
                    for seed in range(S)])
    mean = est.mean(axis=0)
    sem = est.std(axis=0) / math.sqrt(S)
    assert np.all(np.abs(mean - y) < 5 * np.abs(sem) + 1e-12)
    # all targets of one group use the same indices: the estimate of a target
    # equals (N/K) sum over the group's subset
    idx = [oracle.mc_shared_index(3, 0, k, 64) for k in range(K)]
    e = oracle.transform_mc(+1, LAM, amp, phase, u, v, K, 3, shared=True)
    for m in range(3):
        ref = sum(oracle.transform_full(+1, LAM, amp[[n]], phase[[n]], u[[n]], v[[m]])[0]
                  for n in idx)
        assert abs(e[m] - 64 / K * ref) <= 1e-12 * abs(e[m])


def test_mc_indices_uniform(oracle):
    N = 97
    counts = np.zeros(N)
    for m in range(40):
        for k in range(500):
            counts[oracle.mc_index(123, m, k, N)] += 1
    exp = counts.sum() / N
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert chi2 < N + 6 * math.sqrt(2 * N)  # ~6 sigma above the mean of chi2(N-1)


def test_mc_estimator_unbiased(oracle):
    amp, phase, u, v = rand_geom(12, 64, 3)
    y = oracle.transform_full(+1, LAM, amp, phase, u, v)
    K, S = 16, 400
    est = np.stack([oracle.transform_mc(+1, LAM, amp, phase, u, v, K, seed) for seed in range(S)])
    mean = est.mean(axis=0)
    sem = est.std(axis=0) / math.sqrt(S)
    assert np.all(np.abs(mean - y) < 5 * np.abs(sem) + 1e-12)


def test_mc_error_scales_inverse_sqrt_k(oracle):
    amp, phase, u, v = rand_geom(13, 256, 64)
    y = oracle.transform_full(+1, LAM, amp, phase, u, v)
    errs = []
    for K in (16, 64, 256):
        e = [np.linalg.norm(oracle.transform_mc(+1, LAM, amp, phase, u, v, K, s) - y)
             for s in range(12)]
        errs.append(np.mean(e))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 1.5 < r1 < 2.7 and 1.5 < r2 < 2.7  # ~2 for a 4x increase in K


def test_mc_k_equals_all_with_fixed_indices_matches_formula(oracle):
    """The estimator is (N/K) sum over the drawn indices (checked directly)."""
    amp, phase, u, v = rand_geom(14, 30, 2)
    K, seed = 7, 99
    y = oracle.transform_mc(+1, LAM, amp, phase, u, v, K, seed)
    for m in range(2):
        idx = [oracle.mc_index(seed, m, k, 30) for k in range(K)]
        ref = sum(oracle.transform_full(+1, LAM, amp[[n]], phase[[n]], u[[n]], v[[m]])[0]
                  for n in idx)
        assert abs(y[m] - 30 / K * ref) <= 1e-12 * abs(y[m])


# --------------------------------------------------------------------------
# adjoint


def loss_and_field(oracle, amp, theta, u, v, target, sign=+1):
    z = oracle.transform_full(sign, LAM, amp, theta, u, v)
    I = np.abs(z) ** 2
    L = np.mean((I / I.sum() - target) ** 2)
    # dL/dz (as dL/dRe + i dL/dIm)
    S = I.sum()
    dLdI = 2 * (I / S - target) / I.size
    dLdI = (dLdI - np.sum(dLdI * I) / S) / S
    g = 2 * dLdI * z
    return L, g


@pytest.mark.parametrize("sign", [+1, -1])
def test_adjoint_finite_differences(oracle, sign):
    # near field (z = 2 mm): the FP64 rounding of a ~2e4 rad argument
    # (~4e-12 rad) stays far below the step h; at 0.35 m it would not
    amp, theta, u, v = rand_geom(15, 16, 8, z=2e-3, spread=5e-4)
    target = np.random.default_rng(0).random(8)
    target /= target.sum()
    L, g = loss_and_field(oracle, amp, theta, u, v, target, sign)
    ga, gt = oracle.adjoint(sign, LAM, amp, theta, u, g, v)
    h = 1e-5
    for n in range(16):
        tp, tm = theta.copy(), theta.copy()
        tp[n] += h
        tm[n] -= h
        fd = (loss_and_field(oracle, amp, tp, u, v, target, sign)[0]
              - loss_and_field(oracle, amp, tm, u, v, target, sign)[0]) / (2 * h)
        assert abs(fd - gt[n]) <= 1e-5 * max(abs(gt).max(), 1e-30)
        ap, am = amp.copy(), amp.copy()
        ap[n] += h
        am[n] -= h
        fd = (loss_and_field(oracle, ap, theta, u, v, target, sign)[0]
              - loss_and_field(oracle, am, theta, u, v, target, sign)[0]) / (2 * h)
        assert abs(fd - ga[n]) <= 1e-5 * max(abs(ga).max(), 1e-30)


def test_adjoint_dense_matrix(oracle):
    """b = G^H g with G built densely (independent NumPy evaluation)."""
    amp, theta, u, v = rand_geom(16, 40, 24)
    g = np.random.default_rng(1).standard_normal(24) + 1j * np.random.default_rng(2).standard_normal(24)
    d = np.linalg.norm(v[:, None, :] - u[None, :, :], axis=2)
    G = np.exp(1j * 2 * np.pi / LAM * d) / d  # (M, N)
    b = G.conj().T @ g
    ga, gt = oracle.adjoint(+1, LAM, amp, theta, u, g, v)
    e = np.exp(-1j * theta) * b
    assert np.allclose(ga, e.real, rtol=1e-8, atol=1e-8 * np.abs(e).max())
    assert np.allclose(gt, amp * e.imag, rtol=1e-8, atol=1e-8 * np.abs(e).max())
    # forward through the same dense matrix
    z = oracle.transform_full(+1, LAM, amp, theta, u, v)
    assert np.allclose(z, G @ (amp * np.exp(1j * theta)), rtol=1e-8, atol=1e-8 * np.abs(z).max())


# --------------------------------------------------------------------------
# obliquity factor (matlab/holofunctions.m:209-233, lightfieldcalc)


def lightfieldcalc(obs_x, obs_y, holo_angles, holo_field, scr_x, scr_y, z, lam):
    """NumPy restatement of MATLAB lightfieldcalc (holofunctions.m:209-233)
    in double precision: for every screen point k, d = sqrt(dx^2 + dy^2 + z^2),
    cosfact = z / d, sample += cosfact * (field_k / d) * exp(i (2 pi d / lam
    + angle_k)).  Observers at height z above the screen (z > 0)."""
    out = np.zeros(obs_x.size, dtype=np.complex128)
    for k in range(scr_x.size):
        d = np.sqrt((obs_x - scr_x[k]) ** 2 + (obs_y - scr_y[k]) ** 2 + z ** 2)
        out += (z / d) * (holo_field[k] / d) * np.exp(1j * (2 * np.pi * d / lam + holo_angles[k]))
    return out


@pytest.mark.parametrize("sign", [+1, -1])
def test_obliquity_single_source_closed_form(oracle, sign):
    """One source: |y| = a |dz| / d^2 (mpmath), and the phase is that of the
    unweighted pair (the cosine is real and positive); below the source
    plane (dz < 0) the factor stays positive."""
    for v in ([1e-4, -2e-4, 0.05], [3e-4, 1e-4, -0.02], [0.0, 0.0, LAM_REF]):
        re, im, d, arg = oracle.pair(sign, LAM, 0.7, 1.3, [0, 0, 0], v, obliquity=True)
        r0, i0, d0, a0 = oracle.pair(sign, LAM, 0.7, 1.3, [0, 0, 0], v)
        assert d == d0 and arg == a0
        dm = mp.sqrt(mp.mpf(v[0]) ** 2 + mp.mpf(v[1]) ** 2 + mp.mpf(v[2]) ** 2)
        expect = float(mp.mpf("0.7") * abs(mp.mpf(v[2])) / dm ** 2)
        assert math.hypot(re, im) == pytest.approx(expect, rel=4e-16)
        ratio = complex(re, im) / complex(r0, i0)
        assert abs(ratio.imag) <= 1e-15 * abs(ratio) and ratio.real > 0
        assert ratio.real == pytest.approx(float(abs(mp.mpf(v[2])) / dm), rel=4e-16)


def test_obliquity_matches_matlab_lightfieldcalc(oracle):
    """Forwards sum with obliquity == lightfieldcalc (MATLAB sign: + 2 pi d/lam).
    The two formulations round the ~5e5 rad argument differently (2 pi d / lam
    vs d K), about 1e-10 rad, so the bar is 1e-9; a factor of d, a missing
    |.|, or the wrong sign is off by orders of magnitude."""
    g = np.random.default_rng(31)
    n_scr, n_obs, z = 300, 64, 0.05
    sx, sy = (g.random(n_scr) - 0.5) * 2e-3, (g.random(n_scr) - 0.5) * 2e-3
    ox, oy = (g.random(n_obs) - 0.5) * 2e-3, (g.random(n_obs) - 0.5) * 2e-3
    ang, fld = g.random(n_scr) * 2 * np.pi, g.random(n_scr)
    ref = lightfieldcalc(ox, oy, ang, fld, sx, sy, z, LAM)
    u = np.stack([sx, sy, np.zeros(n_scr)], 1)
    v = np.stack([ox, oy, np.full(n_obs, z)], 1)
    y = oracle.transform_full(+1, LAM, fld, ang, u, v, obliquity=True)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-9
    # observers below the screen: |dz| keeps the factor positive
    v[:, 2] = -z
    y = oracle.transform_full(+1, LAM, fld, ang, u, v, obliquity=True)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-9
    # and without the factor the sums differ by about (1 - z/d) ~ 1e-3
    y0 = oracle.transform_full(+1, LAM, fld, ang, u, np.abs(v))
    assert np.linalg.norm(y0 - ref) / np.linalg.norm(ref) > 1e-5


def test_obliquity_cartesian_and_mc_consistent(oracle):
    """Complex weights w = a e^{i phi} give the polar sum (to the argument
    rounding); the MC estimator with obliquity is (N/K) sum over its drawn
    indices of the obliquity pair."""
    amp, phase, u, v = rand_geom(32, 40, 5, z=2e-3, spread=5e-4)
    yp = oracle.transform_full(-1, LAM, amp, phase, u, v, obliquity=True)
    yc = oracle.transform_cartesian(-1, LAM, amp * np.exp(1j * phase), u, v, obliquity=True)
    # fl(phi - d K) vs phi applied after fl(d K): ~4e-12 rad at d ~ 2 mm
    assert np.linalg.norm(yc - yp) / np.linalg.norm(yp) < 1e-10
    K, seed = 9, 5
    y = oracle.transform_mc(+1, LAM, amp, phase, u, v, K, seed, obliquity=True)
    for m in range(5):
        idx = [oracle.mc_index(seed, m, k, 40) for k in range(K)]
        ref = sum(oracle.transform_full(+1, LAM, amp[[n]], phase[[n]], u[[n]], v[[m]],
                                        obliquity=True)[0] for n in idx)
        assert abs(y[m] - 40 / K * ref) <= 1e-12 * abs(y[m])


def test_obliquity_adjoint_dense_matrix(oracle):
    """b = G^H g with G_mn = (|dz|/d) e^{i K d} / d built densely."""
    g0 = np.random.default_rng(33)
    amp, theta, u, v = rand_geom(34, 30, 20, z=0.01, spread=2e-3)
    u[:, 2] = (g0.random(30) - 0.5) * 1e-3  # non-planar sources
    g = g0.standard_normal(20) + 1j * g0.standard_normal(20)
    diff = v[:, None, :] - u[None, :, :]
    d = np.linalg.norm(diff, axis=2)
    G = np.abs(diff[:, :, 2]) / d * np.exp(1j * 2 * np.pi / LAM * d) / d
    ga, gt = oracle.adjoint(+1, LAM, amp, theta, u, g, v, obliquity=True)
    e = np.exp(-1j * theta) * (G.conj().T @ g)
    assert np.allclose(ga, e.real, rtol=1e-8, atol=1e-8 * np.abs(e).max())
    assert np.allclose(gt, amp * e.imag, rtol=1e-8, atol=1e-8 * np.abs(e).max())
    z = oracle.transform_full(+1, LAM, amp, theta, u, v, obliquity=True)
    assert np.allclose(z, G @ (amp * np.exp(1j * theta)), rtol=1e-8, atol=1e-8 * np.abs(z).max())


# --------------------------------------------------------------------------
# the NumPy restatement of the reference's CPU path


def test_numpy_baseline_matches_oracle():
    import oracle as o
    from oracle import numpy_baseline as nb
    amp, phase, u, v = rand_geom(17, 200, 16)
    x = np.stack([amp, phase], axis=1)
    for sign in (+1, -1):
        y = np.empty((16, 2))
        nb.map_sum_kernel(x, y, u, v, direction=-sign, lam=LAM)
        ref = o.transform_full(sign, LAM, amp, phase, u, v)
        assert np.allclose(y[:, 0], np.abs(ref), rtol=1e-8)
        dphi = np.angle(np.exp(1j * (y[:, 1] - np.angle(ref))))
        assert np.all(np.abs(dphi) < 1e-7)


def test_numpy_baseline_known_answer():
    """py/test_util.py:63-64 through the NumPy restatement itself."""
    from oracle import numpy_baseline as nb
    x = np.array([[1.0, 0.0]])
    u = np.zeros((1, 3))
    v = np.array([0.0, 0.0, LAM_REF])
    c = nb.sum_kernel(x, u, v, direction=1, lam=LAM_REF)
    a, phi = nb.from_polar(c)
    assert a == 1538461.5384615385 and phi == 2.4492935982947064e-16
