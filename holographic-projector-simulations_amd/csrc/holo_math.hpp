// FP64 sin/cos and sqrt/reciprocal-sqrt for the superposition inner loop.
//
// hm_sincos: Cody & Waite argument reduction (W. J. Cody, W. Waite,
// "Software Manual for the Elementary Functions", 1980) modulo pi with fused
// multiply-adds and a two-term pi, followed by near-minimax polynomials on
// [-pi/2, pi/2]:
//   k = rint(x / pi);  r = fma(-k, PI1, x);  r = fma(-k, PI2, r)
//   sin r = r + r^3 P(r^2),  cos r = 1 + r^2 Q(r^2)
//   sin x = (-1)^k sin r,    cos x = (-1)^k cos r
// PI1 = RN(pi), PI2 = RN(pi - PI1); the neglected third term is 3e-33, so for
// |x| < 2^50 the reduced argument is within ~1.2e-16 of x - k pi.  P and Q
// (degree 6 in r^2) are least-squares fits on Chebyshev nodes computed by
// gen_sincos_coeffs.py; their absolute error is 1.2e-16 (sin) and 2.2e-15
// (cos) on [-pi/2, pi/2].  Reducing by pi instead of pi/2 leaves a common
// sign (-1)^k and no sin/cos swap: hm_sincos_unsigned returns that sign as a
// bit mask for the caller to fold into an amplitude with one xor.  Only
// multiplies, FMAs, rint and integer ops are used, so host and device builds
// give identical bits.  Superposition arguments stay below 2^27 rad.
#pragma once
#include <hip/hip_runtime.h>

namespace holo {

#define HM_HD __host__ __device__ __forceinline__

HM_HD void hm_sincos_unsigned(double x, double& s, double& c,
                              unsigned long long& sign_mask) {
#pragma clang fp contract(off)
  constexpr double kOneOverPi = 0x1.45f306dc9c883p-2;
  constexpr double kPi1 = 0x1.921fb54442d18p+1;
  constexpr double kPi2 = 0x1.1a62633145c07p-53;
  constexpr double kMagic = 0x1.8p52;  // 1.5 * 2^52: low word = integer
  const double k = __builtin_rint(x * kOneOverPi);
  double r = __builtin_fma(-k, kPi1, x);
  r = __builtin_fma(-k, kPi2, r);
  const unsigned lo = (unsigned)__builtin_bit_cast(unsigned long long, k + kMagic);
  sign_mask = (unsigned long long)(lo << 31) << 32;  // bit 63 = parity of k
  const double u = r * r;
  double p = -0x1.9e8d5c0fb24cdp-41;
  p = __builtin_fma(p, u, 0x1.60e1e537cfcadp-33);
  p = __builtin_fma(p, u, -0x1.ae63326f52860p-26);
  p = __builtin_fma(p, u, 0x1.71de378613a35p-19);
  p = __builtin_fma(p, u, -0x1.a01a01987bec7p-13);
  p = __builtin_fma(p, u, 0x1.1111111106d2cp-7);
  p = __builtin_fma(p, u, -0x1.5555555555412p-3);
  s = __builtin_fma(r * u, p, r);
  double q = -0x1.837194dec695cp-37;
  q = __builtin_fma(q, u, 0x1.1eadefa1f8550p-29);
  q = __builtin_fma(q, u, -0x1.27e3f6aae8f33p-22);
  q = __builtin_fma(q, u, 0x1.a019fd06a0196p-16);
  q = __builtin_fma(q, u, -0x1.6c16c160f5ed5p-10);
  q = __builtin_fma(q, u, 0x1.555555553bbacp-5);
  q = __builtin_fma(q, u, -0x1.ffffffffffafdp-2);
  c = __builtin_fma(q, u, 1.0);
}

// Flip the sign of x where bit 63 of mask is set.
HM_HD double hm_apply_sign(double x, unsigned long long mask) {
  return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) ^ mask);
}

HM_HD void hm_sincos(double x, double& s, double& c) {
  unsigned long long m;
  hm_sincos_unsigned(x, s, c, m);
  s = hm_apply_sign(s, m);
  c = hm_apply_sign(c, m);
}

// d = sqrt(s), correctly rounded, and inv = 1/d to about an ulp, for normal
// s.  Goldschmidt iteration from the hardware reciprocal square root, then
// one Newton correction of d using h ~ 1/(2 sqrt s) (Markstein's final
// step), then one Newton step of the reciprocal 1/d from 2h:
// inv = 2h + 2h (1 - 2 d h).  On gfx950 a second correction changed d on
// none of 4.3e9 random s (in [0.1, 0.2], [1e-6, 1e-2], [0.5, 4] and
// [1e-12, 1e3]), and the two-correction d matched the host's IEEE sqrt on
// 1.7e7 random s in [1e-8, 20]; inv is used only for amplitudes.  On the
// host the seed comes from 1/sqrt(s).
HM_HD void hm_sqrt_inv(double s, double& d, double& inv) {
#pragma clang fp contract(off)
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(s);
#else
  const double y = 1.0 / __builtin_sqrt(s);
#endif
  double g = s * y;
  double h = 0.5 * y;
  double e = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  const double rr = __builtin_fma(-g, g, s);
  g = __builtin_fma(rr, h, g);
  const double h2 = h + h;
  e = __builtin_fma(-g, h2, 1.0);
  d = g;
  inv = __builtin_fma(h2, e, h2);
}

// d = sqrt(s) to within an ulp and inv = 1/d to ~1e-13 relative (the
// Goldschmidt step and one Newton correction of hm_sqrt_inv, without the
// final correction and the reciprocal refinement).  For the FP32-split
// kernels, which need d only for the fractional cycle d / lambda (an ulp of
// d is ~1e-11 cycles at 0.35 m) and 1/d only in FP32.
HM_HD void hm_sqrt_inv_fast(double s, double& d, double& inv) {
#pragma clang fp contract(off)
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(s);
#else
  const double y = 1.0 / __builtin_sqrt(s);
#endif
  double g = s * y;
  double h = 0.5 * y;
  const double e = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  const double rr = __builtin_fma(-g, g, s);
  d = __builtin_fma(rr, h, g);
  inv = h + h;
}

#undef HM_HD

}  // namespace holo
