// Device math for the Huygens superposition on gfx950 (CDNA4, wave64).
//
// pair_f64: the FP64 pair of cuda/superposition.cu:35-51 with the operation
// order pinned (each step one IEEE rounding):
//   dx = v0-u0, dy = v1-u1, dz = v2-u2
//   s   = fma(dz, dz, fma(dy, dy, dx*dx))
//   d   = sqrt(s)                         (correctly rounded)
//   arg = fma(+/-d, K, phi)               (phase +/- distance*K, one rounding)
//   r   = a / d
// The phase argument reaches ~3.5e6 rad at 0.35 m and 632.8 nm, where one ulp
// of d moves the phase by ~5e-10 rad, so d and arg must be bit-identical to
// the CPU restatement for FP64 results to agree to 1e-12 relative L2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "holo_math.hpp"

namespace holo {

constexpr int kBlock = 256;  // 4 wave64 per workgroup

struct PairF64 {
  double r;    // amplitude a/d
  double arg;  // phase argument
};

// MATH = 0: hm_sqrt_hinv (d correctly rounded, 1/d within 3.7e-15); MATH = 1: the
// compiler's sqrt and IEEE division.  Both give the same d and arg bits.
// OBL: amplitude times the obliquity factor |dz| / d
// (matlab/holofunctions.m:226-227), r = (a / d) (|dz| / d).
// pair_f64_s takes dxx = dx * dx (rounded), dy and dz already formed (dz once
// per target when every source has the same z, dxx once per row of sources
// with one x); pair_f64_d takes dx = vx - ux; pair_f64 forms them all.
template <int SIGN, int MATH = 0, bool OBL = false>
__device__ __forceinline__ PairF64 pair_f64_s(double a, double phi, double dxx, double dy,
                                              double dz, double K) {
#pragma clang fp contract(off)
  const double s = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dxx));
  PairF64 p;
  if constexpr (MATH == 0) {
    double d, h;  // h = (1/d) / 2
    hm_sqrt_hinv(s, d, h);
    p.arg = __builtin_fma(SIGN > 0 ? d : -d, K, phi);
    p.r = (a + a) * h;  // = a (h + h) exactly; a + a is shared by a lane's targets
    if constexpr (OBL) p.r *= __builtin_fabs(dz) * (h + h);
  } else {
    const double d = __builtin_sqrt(s);
    p.arg = __builtin_fma(SIGN > 0 ? d : -d, K, phi);
    p.r = a / d;
    if constexpr (OBL) p.r *= __builtin_fabs(dz) / d;
  }
  return p;
}

template <int SIGN, int MATH = 0, bool OBL = false>
__device__ __forceinline__ PairF64 pair_f64_d(double a, double phi, double dx, double dy,
                                              double dz, double K) {
#pragma clang fp contract(off)
  return pair_f64_s<SIGN, MATH, OBL>(a, phi, dx * dx, dy, dz, K);
}

template <int SIGN, int MATH = 0, bool OBL = false>
__device__ __forceinline__ PairF64 pair_f64(double a, double phi, double ux,
                                            double uy, double uz, double vx,
                                            double vy, double vz, double K) {
#pragma clang fp contract(off)
  return pair_f64_d<SIGN, MATH, OBL>(a, phi, vx - ux, vy - uy, vz - uz, K);
}

// r * (cos x, sin x) accumulated into (ar, ai): hm_sincos with the table at
// tab (MATH = 0) or the device libm sincos (MATH = 1).
template <int MATH = 0>
__device__ __forceinline__ void accumulate_phasor(double r, double x, double& ar, double& ai,
                                                  const double* tab = kSinCosTab) {
  double s, c;
  if constexpr (MATH == 0) {
    hm_sincos(x, tab, s, c);
  } else {
    sincos(x, &s, &c);
  }
  ar = __builtin_fma(r, c, ar);
  ai = __builtin_fma(r, s, ai);
}

// Distance only (same pinned order as pair_f64).
__device__ __forceinline__ double distance_f64(double ux, double uy, double uz,
                                               double vx, double vy,
                                               double vz) {
#pragma clang fp contract(off)
  const double dx = vx - ux;
  const double dy = vy - uy;
  const double dz = vz - uz;
  return __builtin_sqrt(__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)));
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11), counter-based: no state arrays.
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
  }
  return c;
}

constexpr uint32_t kMcStream = 0x4d43u;  // third counter word of MC draws

// Lemire multiply-shift of the 64-bit uniform word r = hi 2^32 + lo into
// [0, N): floor(r N / 2^64).  With N < 2^32 that is
// floor((hi N + floor(lo N / 2^32)) / 2^32) exactly (the dropped fraction is
// below one and the sum below 2^64), two 32 x 32 -> 64-bit products.
__device__ __forceinline__ uint32_t lemire_index(uint32_t lo, uint32_t hi, uint32_t N) {
  const uint64_t t = (uint64_t)hi * N + (((uint64_t)lo * N) >> 32);
  return (uint32_t)(t >> 32);
}

// Fractional cycle in [-0.5, 0.5]: x - rint(x), exact in FP64 for |x| < 2^52.
__device__ __forceinline__ double frac_cycle(double x) {
  return x - __builtin_rint(x);
}

// sin/cos of 2*pi*t for t in revolutions (v_sin_f32 / v_cos_f32 take
// revolutions).
__device__ __forceinline__ void sincos_rev_f32(float t, float& s, float& c) {
  s = __builtin_amdgcn_sinf(t);
  c = __builtin_amdgcn_cosf(t);
}

// ---------------------------------------------------------------------------
// FP32-split helpers shared by the tiled and lattice kernels.

// One pair of the FP32-split formulation from its squared distance s and
// axial offset dz: d (to a few 1e-15, not always correctly rounded) and the
// fractional cycle in FP64, sin/cos in FP32, FP32 accumulation (flushed to
// FP64 by the caller).  OBL: amplitude |dz| / d^2 instead of 1 / d.
template <int SIGN, bool OBL = false>
__device__ __forceinline__ void pair_f32s_s(double s, double dz, float2 w, double inv_lambda,
                                            float& fr, float& fi) {
  double d, id;
  hm_sqrt_inv_fast(s, d, id);
  // the cycle's fraction in [0, 1), exact (v_fract_f64), then rounded to FP32
  // (within 3e-8 cycles)
  const float f = (float)__builtin_amdgcn_fract(d * inv_lambda);
  float sn, cs;
  sincos_rev_f32(f, sn, cs);
  if (SIGN < 0) sn = -sn;
  const float idf = OBL ? (float)(id * (__builtin_fabs(dz) * id)) : (float)id;
  const float pr = idf * cs, pi = idf * sn;
  fr = __builtin_fmaf(w.x, pr, __builtin_fmaf(-w.y, pi, fr));
  fi = __builtin_fmaf(w.x, pi, __builtin_fmaf(w.y, pr, fi));
}

// One sample of the FP32-split estimator at the source p (position in
// .x .y .z, the FP32 weight packed in .w).
template <int SIGN, bool OBL = false>
__device__ __forceinline__ void mc_pair_f32s(const double4& p, double vx, double vy, double vz,
                                             double inv_lambda, float& fr, float& fi) {
  const double dx = vx - p.x, dy = vy - p.y, dz = vz - p.z;
  // the FP32 weight packed in .w, unpacked by integer ops (no trip through memory)
  const uint64_t wb = __builtin_bit_cast(uint64_t, p.w);
  const float2 w = make_float2(__builtin_bit_cast(float, (uint32_t)wb),
                               __builtin_bit_cast(float, (uint32_t)(wb >> 32)));
  pair_f32s_s<SIGN, OBL>(__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)), dz, w,
                         inv_lambda, fr, fi);
}

// Accuracy guard: per (target, tile) the bound q_max of |q| over the tile
// (from the tile's lateral and axial radii) gives the FP32 rounding error of
// t, ~eps32 q_max / (2 d lambda) cycles, and the truncation of the q-series,
// q_max^3 / (16 d^5 lambda) cycles.  If either exceeds kGuardCycles (thick
// 3-D tiles, near field) that target evaluates the tile exactly: FP64
// distance from the source's FP64 position, FP64 fractional cycle, FP32
// sin/cos and weights.
constexpr float kGuardCycles = 2e-7f;

// One exact pair of the FP32-split formulation at the source (ux, uy, uz):
// d^2 = dy^2 + (dz^2 + dx^2), in that order, so that a lattice row's
// dz^2 + dx^2 can be formed once (exact_lattice_tile) with the same bits.
template <int SIGN, bool OBL>
__device__ __forceinline__ void exact_pair(double ux, double uy, double uz, float2 w, double vx,
                                           double vy, double vz, double inv_lambda, float& fr,
                                           float& fi) {
  const double dx = vx - ux, dy = vy - uy, dz = vz - uz;
  pair_f32s_s<SIGN, OBL>(__builtin_fma(dy, dy, __builtin_fma(dz, dz, dx * dx)), dz, w,
                         inv_lambda, fr, fi);
}

// exact_tile_pairs and exact_lattice_tile sum their pairs in FP32 over runs of
// kExactRun slots (a lattice tile's row), each run's sum then added in FP64:
// over a 16 x 16 tile that cuts the FP32 summation error about 3x (the
// tiled kernel's tiles of at most 64 are summed whole).
constexpr int kExactRun = 16;

// The exact sum of one target over the T slots of a tile (weights as
// (re, im); slot j holds source perm[j]): exact_pair per source, in slot
// order, in runs of kExactRun slots.  Sources come 8 at a time, their
// (wave-uniform) loads issued before any of them is evaluated, so that the
// dependent loads through perm overlap; padding slots (perm < 0) are skipped.
template <int SIGN, bool OBL>
__device__ __forceinline__ void exact_tile_pairs(
    const double* __restrict__ u, const int32_t* __restrict__ perm,
    const float2* __restrict__ w, int T, double vx, double vy, double vz,
    double inv_lambda, double& ar, double& ai) {
  float fr = 0.f, fi = 0.f;
  int j0 = 0;
  for (; j0 + 8 <= T; j0 += 8) {
    int32_t n[8];
    float2 wj[8];
    double4 p[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) n[h] = perm[j0 + h];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      wj[h] = w[j0 + h];
      const size_t m = n[h] < 0 ? 0 : (size_t)n[h];  // a padding slot reads source 0, unused
      p[h] = make_double4(u[3 * m], u[3 * m + 1], u[3 * m + 2], 0.0);
    }
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      if (n[h] < 0) continue;
      exact_pair<SIGN, OBL>(p[h].x, p[h].y, p[h].z, wj[h], vx, vy, vz, inv_lambda, fr, fi);
    }
    if ((j0 + 8) % kExactRun == 0) {
      ar += fr;
      ai += fi;
      fr = fi = 0.f;
    }
  }
  for (int j = j0; j < T; ++j) {
    const int32_t n = perm[j];
    if (n >= 0)
      exact_pair<SIGN, OBL>(u[3 * (size_t)n], u[3 * (size_t)n + 1], u[3 * (size_t)n + 2], w[j],
                            vx, vy, vz, inv_lambda, fr, fi);
    if ((j + 1) % kExactRun == 0) {
      ar += fr;
      ai += fi;
      fr = fi = 0.f;
    }
  }
  ar += fr;
  ai += fi;
}

// One exact pair from its squared distance s and axial offset dz, with the
// weight in polar form, w = (|w|, arg w in cycles): the weight's phase joins
// the pair's fractional cycle before sin/cos and its modulus the amplitude
// (FP64 distance and cycle fraction as in pair_f32s_s, the sum of the two
// fractions in FP32, within 6e-8 cycles).
template <int SIGN, bool OBL>
__device__ __forceinline__ void polar_pair_s(double s, double dz, float2 w, double inv_lambda,
                                             float& fr, float& fi) {
  double d, id;
  hm_sqrt_inv_fast(s, d, id);
  const float f = (float)__builtin_amdgcn_fract(d * inv_lambda);
  float sn, cs;
  sincos_rev_f32(SIGN > 0 ? w.y + f : w.y - f, sn, cs);
  const float idf = OBL ? (float)(id * (__builtin_fabs(dz) * id)) : (float)id;
  const float amp = w.x * idf;
  fr = __builtin_fmaf(amp, cs, fr);
  fi = __builtin_fmaf(amp, sn, fi);
}

// The exact pairs of one tile for those of a lane's R targets whose flag
// near[r] is set, with polar weights (the tiled kernel's).  Each source is
// loaded once for all R targets, 8 sources at a time as in exact_tile_pairs;
// padding slots are skipped.  Every target's pairs come in slot order.
// PLANAR: every source of the tile has the z of its centre, cz, so that dz
// and dz^2 are formed once per target and d^2 = dy^2 + (dx^2 + dz^2).
template <int SIGN, bool OBL, int R, bool PLANAR>
__device__ __forceinline__ void exact_tile_pairs_polar(
    const double* __restrict__ u, const int32_t* __restrict__ perm,
    const float2* __restrict__ w, int T, double cz, const double (&vx)[R],
    const double (&vy)[R], const double (&vz)[R], const bool (&near)[R], double inv_lambda,
    double (&ar)[R], double (&ai)[R]) {
  float fr[R], fi[R];
  double dzp[R], dz2[R];  // PLANAR: vz - cz and its square
#pragma unroll
  for (int r = 0; r < R; ++r) {
    fr[r] = fi[r] = 0.f;
    dzp[r] = vz[r] - cz;
    dz2[r] = dzp[r] * dzp[r];
  }
  const auto pair = [&](int r, double px, double py, double pz, float2 wj)
                        __attribute__((always_inline)) {
    const double dx = vx[r] - px, dy = vy[r] - py;
    if constexpr (PLANAR) {
      polar_pair_s<SIGN, OBL>(__builtin_fma(dy, dy, __builtin_fma(dx, dx, dz2[r])), dzp[r], wj,
                              inv_lambda, fr[r], fi[r]);
    } else {
      const double dz = vz[r] - pz;
      polar_pair_s<SIGN, OBL>(__builtin_fma(dy, dy, __builtin_fma(dz, dz, dx * dx)), dz, wj,
                              inv_lambda, fr[r], fi[r]);
    }
  };
  int j0 = 0;
  for (; j0 + 8 <= T; j0 += 8) {
    int32_t n[8];
    float2 wj[8];
    double4 p[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) n[h] = perm[j0 + h];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      wj[h] = w[j0 + h];
      const size_t m = n[h] < 0 ? 0 : (size_t)n[h];  // a padding slot reads source 0, unused
      p[h] = make_double4(u[3 * m], u[3 * m + 1], PLANAR ? 0.0 : u[3 * m + 2], 0.0);
    }
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      if (n[h] < 0) continue;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (near[r]) pair(r, p[h].x, p[h].y, p[h].z, wj[h]);
    }
  }
  for (int j = j0; j < T; ++j) {
    const int32_t n = perm[j];
    if (n < 0) continue;
    const double px = u[3 * (size_t)n], py = u[3 * (size_t)n + 1];
    const double pz = PLANAR ? 0.0 : u[3 * (size_t)n + 2];
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (near[r]) pair(r, px, py, pz, w[j]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (near[r]) {
      ar[r] += fr[r];
      ai[r] += fi[r];
    }
}

// The exact sum of one target over one 16 x 16 tile of a uniform lattice,
// whose rows have one x and columns one y each: xs holds the tile's 16 row x's,
// ys its 16 column y's (entries beyond the lattice finite and unused), z the
// lattice's z, w the tile's 256 weights (slot 16 i + j = row i, column j).
// Rows below ni and columns below nj are sources, the other slots padding.
// dz^2 + dx^2 is formed once per row and dy once per column; each pair then
// takes the operations of exact_pair, in the same order, the pairs come in
// slot order and each row is one run of kExactRun slots, so the sum has the
// bits of exact_tile_pairs over the tile's slots, without the per-pair
// gathers of three positions through perm.
template <int SIGN, bool OBL>
__device__ __forceinline__ void exact_lattice_tile(
    const double* __restrict__ xs, const double* __restrict__ ys, double z, int ni, int nj,
    const float2* __restrict__ w, double vx, double vy, double vz, double inv_lambda,
    double& ar, double& ai) {
  float fr = 0.f, fi = 0.f;
  const double dz = vz - z;
  double dy[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) dy[j] = vy - ys[j];
  if (nj == 16) {  // every column a source: straight-line rows
    for (int i = 0; i < ni; ++i) {
      const double dx = vx - xs[i];
      const double a = __builtin_fma(dz, dz, dx * dx);
#pragma unroll
      for (int j = 0; j < 16; ++j)
        pair_f32s_s<SIGN, OBL>(__builtin_fma(dy[j], dy[j], a), dz, w[16 * i + j], inv_lambda,
                               fr, fi);
      ar += fr;
      ai += fi;
      fr = fi = 0.f;
    }
  } else {
    for (int i = 0; i < ni; ++i) {
      const double dx = vx - xs[i];
      const double a = __builtin_fma(dz, dz, dx * dx);
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j < nj)
          pair_f32s_s<SIGN, OBL>(__builtin_fma(dy[j], dy[j], a), dz, w[16 * i + j], inv_lambda,
                                 fr, fi);
      ar += fr;
      ai += fi;
      fr = fi = 0.f;
    }
  }
  ar += fr;
  ai += fi;
}

}  // namespace holo
