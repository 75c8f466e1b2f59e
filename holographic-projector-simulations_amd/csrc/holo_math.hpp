// FP64 sin/cos and sqrt/reciprocal-sqrt for the superposition inner loop.
//
// hm_sincos: Cody & Waite argument reduction (W. J. Cody, W. Waite,
// "Software Manual for the Elementary Functions", 1980) modulo pi/1024 with
// fused multiply-adds and a two-term pi/1024, a table of (sin, cos)(j pi/1024)
// and the angle-addition formulas (the table-driven scheme of P. T. P. Tang,
// "Table-driven implementation of the exponential function", 1989, applied
// to sin/cos):
//   k = RN_int(x 1024/pi);  r = fma(-k, P1, x);  r = fma(-k, P2, r);  j = k mod 2048
//   sin x = S_j cos r + C_j sin r,   cos x = C_j cos r - S_j sin r
// P1 = RN(pi/1024), P2 = RN(pi/1024 - P1); the neglected third term is below
// 2e-35, so for |x| < 2^40 the reduced argument is within ~2e-19 of
// x - k pi/1024.  |r| <= pi/2048, where sin r = r - r^3/6 is within 7.1e-17
// and cos r = 1 + C2 r^2 within 3.96e-14: C2 = -1/2 + (sqrt 2 - 1) R^2 / 12,
// R = pi/2048, is the minimax choice for that one term (the error
// r^4/24 - (C2 + 1/2) r^2 takes +-3.96e-14 at r = R and r^2 = 6 (C2 + 1/2);
// the Taylor term -1/2 alone would leave up to 2.3e-13, and the r^4/24 term
// costs one more FMA per pair, 3 % of the FP64 kernel's time).  The error e
// moves (sin x, cos x) by e (S_j, C_j), along the part (S_j, C_j) cos r of the
// result: an amplitude error below 4e-14 and no phase error, against the
// 1e-12 relative L2 to which the FP64 path is held.  The table
// (holo_sincos_table.inc, 32 KB, written by gen_sincos_table.py) is correctly
// rounded, so S_{2048-j} = -S_j and C_{2048-j} = C_j exactly and sin stays
// odd and cos even bit for bit.  Only multiplies, FMAs, rint and integer ops are
// used, so host and device builds give identical bits.  Superposition
// arguments stay below 2^27 rad.
#pragma once
#include <hip/hip_runtime.h>

namespace holo {

#define HM_HD __host__ __device__ __forceinline__

constexpr double kSinCosTab[4096] = {
#include "holo_sincos_table.inc"
};

// sin and cos of x; tab is kSinCosTab or a copy of it (in LDS in the FP64
// kernel).
HM_HD void hm_sincos(double x, const double* tab, double& s, double& c) {
#pragma clang fp contract(off)
  constexpr double kStepsOverPi = 0x1.45f306dc9c883p+8;  // 1024 / pi
  constexpr double kP1 = 0x1.921fb54442d18p-9;
  constexpr double kP2 = 0x1.1a62633145c07p-63;
  // kk = x 1024/pi + 1.5 * 2^52 in one rounding: its ulp is 1, so kk - 1.5 * 2^52
  // is x 1024/pi rounded to the nearest integer k (ties to even, so k(-x) =
  // -k(x)), and the low word of kk is k mod 2^32 (|k| < 2^51).
  constexpr double kShift = 0x1.8p52;
  const double kk = __builtin_fma(x, kStepsOverPi, kShift);
  const double k = kk - kShift;
  double r = __builtin_fma(-k, kP1, x);
  r = __builtin_fma(-k, kP2, r);
  const unsigned off = ((unsigned)__builtin_bit_cast(unsigned long long, kk) << 4) & 0x7ff0u;
  const double* e = reinterpret_cast<const double*>(reinterpret_cast<const char*>(tab) + off);
  const double S = e[0], C = e[1];
  const double u = r * r;
  const double sr = __builtin_fma(r * u, -0x1.5555555555555p-3, r);
  const double cr = __builtin_fma(u, -0x1.fffffa8c96463p-2, 1.0);  // 1 + C2 r^2
  s = __builtin_fma(S, cr, C * sr);
  c = __builtin_fma(C, cr, -(S * sr));
}

HM_HD void hm_sincos(double x, double& s, double& c) { hm_sincos(x, kSinCosTab, s, c); }

// d = sqrt(s), correctly rounded, and inv ~ 1/d, for normal s.  One
// Goldschmidt iteration from the hardware reciprocal square root gives
// g ~ sqrt(s) and h ~ 1/(2 sqrt s); one Newton correction of d with h
// (Markstein's final step) rounds d correctly, and inv = 2h.  On gfx950 a
// second correction changed d on none of 4.3e9 random s (in [0.1, 0.2],
// [1e-6, 1e-2], [0.5, 4] and [1e-12, 1e3]), and the two-correction d
// matched the host's IEEE sqrt on 1.7e7 random s in [1e-8, 20].  inv scales
// amplitudes only; its error against 1/d is bounded on the device by
// tests/test_gpu_parity.py::test_fp64_inverse_distance_accuracy.  On the host
// the seed comes from 1/sqrt(s).  hm_sqrt_hinv returns h = inv / 2 (exactly)
// for callers that fold the factor 2 into a value they compute once.
HM_HD void hm_sqrt_hinv(double s, double& d, double& h) {
#pragma clang fp contract(off)
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(s);
#else
  const double y = 1.0 / __builtin_sqrt(s);
#endif
  double g = s * y;
  h = 0.5 * y;
  const double e = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  const double rr = __builtin_fma(-g, g, s);
  d = __builtin_fma(rr, h, g);
}

HM_HD void hm_sqrt_inv(double s, double& d, double& inv) {
  double h;
  hm_sqrt_hinv(s, d, h);
  inv = h + h;
}

// d ~ sqrt(s) and inv ~ 1/d from the Goldschmidt step alone (no Markstein
// correction): both within a few 1e-15 relative, as inv above, but d is not
// always correctly rounded.  For FP32-split pairs, which need d / lambda to
// ~1e-8 cycles (at 1 m and 632.8 nm, 4e-15 of d is 6e-9 cycles), at 5 FP64
// operations and one v_rsq_f64 instead of 8.
HM_HD void hm_sqrt_inv_fast(double s, double& d, double& inv) {
#pragma clang fp contract(off)
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(s);
#else
  const double y = 1.0 / __builtin_sqrt(s);
#endif
  const double g = s * y;
  const double e = __builtin_fma(-g, 0.5 * y, 0.5);
  d = __builtin_fma(g, e, g);
  inv = __builtin_fma(y, e, y);
}

#undef HM_HD

}  // namespace holo
