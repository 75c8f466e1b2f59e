// FP64 sin/cos and sqrt/reciprocal-sqrt for the superposition inner loop.
//
// hm_sincos: Cody & Waite argument reduction (W. J. Cody, W. Waite,
// "Software Manual for the Elementary Functions", 1980) modulo pi/1024 with
// fused multiply-adds and a two-term pi/1024, a table of (sin, cos)(j pi/1024)
// and the angle-addition formulas (the table-driven scheme of P. T. P. Tang,
// "Table-driven implementation of the exponential function", 1989, applied
// to sin/cos):
//   k = rint(x 1024/pi);  r = fma(-k, P1, x);  r = fma(-k, P2, r);  j = k mod 2048
//   sin x = S_j cos r + C_j sin r,   cos x = C_j cos r - S_j sin r
// P1 = RN(pi/1024), P2 = RN(pi/1024 - P1); the neglected third term is below
// 2e-35, so for |x| < 2^40 the reduced argument is within ~2e-19 of
// x - k pi/1024.  |r| <= pi/2048, where sin r = r - r^3/6 and cos r = 1 -
// r^2/2 + r^4/24 are within 7.1e-17 and 1.8e-20.  The table
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
  // k + 1.5 * 2^48 has ulp 1/16, so its low word is 16 k mod 2^32: the byte
  // offset of entry k mod 2048 is one AND away (|k| < 2^47)
  constexpr double kMagic = 0x1.8p48;
  const double k = __builtin_rint(x * kStepsOverPi);
  double r = __builtin_fma(-k, kP1, x);
  r = __builtin_fma(-k, kP2, r);
  const unsigned off = (unsigned)__builtin_bit_cast(unsigned long long, k + kMagic) & 0x7ff0u;
  const double* e = reinterpret_cast<const double*>(reinterpret_cast<const char*>(tab) + off);
  const double S = e[0], C = e[1];
  const double u = r * r;
  const double sr = __builtin_fma(r * u, -0x1.5555555555555p-3, r);
  const double cr = __builtin_fma(u, __builtin_fma(u, 0x1.5555555555555p-5, -0.5), 1.0);
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

#undef HM_HD

}  // namespace holo
