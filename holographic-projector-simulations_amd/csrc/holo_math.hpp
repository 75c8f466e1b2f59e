// FP64 sin/cos and sqrt/reciprocal-sqrt for the superposition inner loop.
//
// hm_sincos: Cody & Waite argument reduction (W. J. Cody, W. Waite,
// "Software Manual for the Elementary Functions", 1980) with fused
// multiply-adds and a two-term pi/2, followed by Taylor polynomials on
// [-pi/4, pi/4]:
//   k = rint(x * 2/pi);  r = fma(-k, P1, x);  r = fma(-k, P2, r)
//   sin r = r + r^3 (S1 + r^2 (S2 + ... S7)),  S_j = (-1)^j / (2j+1)!
//   cos r = 1 + r^2 (C1 + r^2 (C2 + ... C8)),  C_j = (-1)^j / (2j)!
// P1 = RN(pi/2), P2 = RN(pi/2 - P1); the neglected third term is < 2^-109, so
// for |x| < 2^50 the reduced argument is within ~1.2e-16 of x - k pi/2 and
// the result within ~3e-16 of sin x, cos x.  The polynomial truncation is
// below 7e-17 on [-pi/4, pi/4].  Only multiplies, FMAs and rint are used, so
// the host and device builds produce identical bits.  The superposition's
// arguments stay below 2^27 rad (d < 13 m at 632.8 nm), far inside the domain.
#pragma once
#include <hip/hip_runtime.h>

namespace holo {

#define HM_HD __host__ __device__ __forceinline__

HM_HD void hm_sincos(double x, double& s, double& c) {
#pragma clang fp contract(off)
  constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
  constexpr double kP1 = 0x1.921fb54442d18p+0;
  constexpr double kP2 = 0x1.1a62633145c07p-54;
  constexpr double kMagic = 0x1.8p52;  // 1.5 * 2^52: low word = integer
  const double k = __builtin_rint(x * kTwoOverPi);
  double r = __builtin_fma(-k, kP1, x);
  r = __builtin_fma(-k, kP2, r);
  const double kb = k + kMagic;
  const unsigned q = (unsigned)__builtin_bit_cast(unsigned long long, kb) & 3u;
  const double r2 = r * r;
  double ps = -0x1.ae7f3e733b81fp-41;               // S7
  ps = __builtin_fma(ps, r2, 0x1.6124613a86d09p-33);  // S6
  ps = __builtin_fma(ps, r2, -0x1.ae64567f544e4p-26); // S5
  ps = __builtin_fma(ps, r2, 0x1.71de3a556c734p-19);  // S4
  ps = __builtin_fma(ps, r2, -0x1.a01a01a01a01ap-13); // S3
  ps = __builtin_fma(ps, r2, 0x1.1111111111111p-7);   // S2
  ps = __builtin_fma(ps, r2, -0x1.5555555555555p-3);  // S1
  const double sr = __builtin_fma(r * r2, ps, r);
  double pc = 0x1.ae7f3e733b81fp-45;                  // C8
  pc = __builtin_fma(pc, r2, -0x1.93974a8c07c9dp-37); // C7
  pc = __builtin_fma(pc, r2, 0x1.1eed8eff8d898p-29);  // C6
  pc = __builtin_fma(pc, r2, -0x1.27e4fb7789f5cp-22); // C5
  pc = __builtin_fma(pc, r2, 0x1.a01a01a01a01ap-16);  // C4
  pc = __builtin_fma(pc, r2, -0x1.6c16c16c16c17p-10); // C3
  pc = __builtin_fma(pc, r2, 0x1.5555555555555p-5);   // C2
  pc = __builtin_fma(pc, r2, -0.5);                   // C1
  const double cr = __builtin_fma(pc, r2, 1.0);
  double so = (q & 1u) ? cr : sr;
  double co = (q & 1u) ? sr : cr;
  s = (q & 2u) ? -so : so;
  c = ((q + 1u) & 2u) ? -co : co;
}

// d = sqrt(s), correctly rounded, and inv = 1/d to ~2e-15 relative, for
// normal s.  Goldschmidt iteration from the hardware reciprocal square root,
// then two Newton corrections of d using h ~ 1/(2 sqrt s) (Markstein's final
// step) and one more refinement of h.  On gfx950, d matched the host's IEEE
// sqrt on 1.7e7 random s in [1e-8, 20]; inv is used only for amplitudes.
// On the host the seed comes from 1/sqrt(s).
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
  double rr = __builtin_fma(-g, g, s);
  g = __builtin_fma(rr, h, g);
  rr = __builtin_fma(-g, g, s);
  g = __builtin_fma(rr, h, g);
  e = __builtin_fma(-g, h, 0.5);
  h = __builtin_fma(h, e, h);
  d = g;
  inv = h + h;
}

#undef HM_HD

}  // namespace holo
