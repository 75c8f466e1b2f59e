// Lattice kernel: FP32-split superposition for sources on a product grid
// (every plane of the reference's generators, cuda/init.h:89-158), with the
// sum over each 8 x 8 block of sources evaluated as a small complex GEMM on
// the matrix cores.
//
// For target m and tile t (centre c, r = v_m - c, d_ref = |r|), the source at
// offset (gx_a, gy_b, 0) has q = d^2 - d_ref^2 = P_a + Q_b with
//   P_a = gx_a^2 - 2 r_x gx_a,    Q_b = gy_b^2 - 2 r_y gy_b.
// The tiled FP32-split expansion (d - d_ref)/lambda = alpha q + c2 q^2 and
// 1/d = (1/d_ref)(1 + kappa q) then factorises into a row factor, a column
// factor and a small cross term:
//   e^{i s 2 pi d/lambda} / d ~= e^{i s 2 pi f} X_a Y_b (1 + i s beta P_a Q_b)
//   X_a = (1 + kappa P_a) e^{i s 2 pi (alpha P_a + c2 P_a^2)}
//   Y_b = (1/d_ref)(1 + kappa Q_b) e^{i s 2 pi (alpha Q_b + c2 Q_b^2)},  beta = 4 pi c2
// with f = frac(d_ref / lambda) from FP64.  Dropped: kappa^2 P Q in the
// amplitude and (beta P Q)^2 / 2 in the phasor; the accuracy guard bounds
// both, with the tiled kernel's rounding and truncation bounds, and sends a
// (target, tile) that fails any of them through exact FP64-distance pairs.
// In the cross term, P_a = gx_a^2 - 2 r_x gx_a is dominated by its linear
// part (the gx^2 share of beta P Q is below 1e-7 rad), so with the linear
// part only a tile contributes
//   S = sum_a X_a T_a,   T_a = sum_b W_ab Y_b + sum_b (i gx_a W_ab) (-2 s beta r_x Q_b Y_b)
// and T for 16 targets at once is a 16 x 16 real product (complex 8 x 8
// written out) on the matrix cores.  Both operands are split into f16 pairs,
// x = x_hi + x_lo (W rounded to nearest, |W_lo| <= 2^-11 |W|; Y rounded
// toward zero, |Y_lo| < 2^-10 |Y|), and two v_mfma_f32_16x16x32_f16
// accumulate, in FP32,
//   W_hi Y_hi + W_hi Y_lo   and   W_lo Y_hi + (i gx W) (correction)
// leaving out only W_lo Y_lo (below 2^-21 of each term, of either sign); the correction is
// below 1e-4 of T, so f16 suffices for it.  The weights are divided by a power
// of two tscale per tile and Y is kept near unit size (Y' = Y / A), so the
// f16 range holds; X carries tscale and the amplitude A back.
//
// Work split (wave of 64 lanes, 64 targets): lane l = 16 g + j computes the
// FP64 constants of target l for the tile and shares them through LDS; for
// each block B of 16 targets it then supplies Y of columns 2g, 2g+1 of target
// 16 B + j as the GEMM's B operand and receives Z of rows 2g, 2g+1, which it
// multiplies by its X.  Operand layout of 16x16x32 f16: lane l holds eight
// consecutive k of row l%16 of A and of column l%16 of B, and D[4(l/16)+r][l%16];
// A and B use the same k for the same (lane group g, slot e), and the slots are
//   e < 4: column b = 2g + e/2, component e%2 (re, im), hi part
//   e >= 4: the same column and component, lo part or correction
//   i = 2a + c  <->  row a, component c
// so the weights enter as A[2a+c][2b+c'] = (Wr, -Wi; Wi, Wr) (k_grid_weights).
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

#include "holo_device.hpp"
#include "holo_internal.hpp"

namespace holo {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));

constexpr int kGridWaves = kBlock / 64;
constexpr int kGridFlush = 8;  // tiles summed in FP32 between FP64 flushes
constexpr float kFourPi = 12.566370614359172f;

// Lattice-wide bounds for the once-per-launch accuracy check.
struct GridBounds {
  double cx0, cx1, cy0, cy1, z;
  float rho_x, rho_y;
  int series = 0;  // 16-row kernels: near pairs take the series path where it holds
  int near_gemm = 0;  // 16-row kernels: waves off the far-field bounds take the near-field GEMM
};

// A uniform lattice's 16 x 16 tiles for the exact path, beside the table
// `axes` of their row x's (16 per row of tiles) followed, from index ys, by
// their column y's (16 per column of tiles): the number of tile columns and
// the lattice's sides.  (The table is a kernel argument of its own, so that
// its loads, wave-uniform, go through the scalar cache.)
struct LatticeAxes {
  uint32_t ys, tj, nx, ny;
};

// The accuracy bounds of one (target, tile), or of a target against every
// tile when given the largest lateral offsets and the smallest distance: FP32
// rounding of the phase, truncation of the q-series, the dropped second order
// of the cross term, the dropped gx^2 share of the cross term (rx2 = largest
// gx^2) and the dropped second order of the amplitude factorisation.  Written
// so that NaN (a target at a tile centre) fails.
__device__ __forceinline__ bool grid_pairs_ok(float Pm, float Qm, float rx2, float ia,
                                              float idf, float kappa) {
  const float qmax = Pm + Qm;
  const float round_err = 3e-8f * qmax * ia;
  const float trunc_err = 0.0625f * qmax * qmax * qmax * ia * idf * idf * idf * idf;
  const float beta = kFourPi * 0.125f * ia * idf * idf;  // |beta| = 4 pi |c2|
  const float cross = beta * Pm * Qm;                    // radians
  const float ampsep = kappa * kappa * Pm * Qm;
  return round_err <= kGuardCycles && trunc_err <= 0.05f * kGuardCycles && cross <= 1.5e-3f &&
         beta * rx2 * Qm <= 1e-6f && ampsep <= 1e-6f;
}

// x = hi + lo in f16: hi is x rounded to 11 significant bits (ties away from
// zero) by integer arithmetic, so hi is exact in f16 for |x| >= 2^-14 and
// every conversion instruction yields the same hi; lo = x - hi is exact in
// FP32 and |lo| <= 2^-11 |x|.  (Splitting as lo = x - (float)(half)x is not
// safe: the compiler may fold the product that made x into the f16
// conversion -- one rounding instead of two -- for one use of hi and not for
// another, and at exact ties hi + lo is then off by one f16 ulp.)
__device__ __forceinline__ void split_f16(float x, _Float16& hi, _Float16& lo) {
  const float h =
      __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, x) + 0x1000u) & 0xffffe000u);
  hi = (_Float16)h;
  lo = (_Float16)(x - h);
}

// The GEMM's Y operand for one column, Y' = am (c, s) with am = 1 + kappa Q
// within ~1e-5 of 1: hi = (c, s) rounded toward zero to f16 by one
// v_cvt_pkrtz_f16_f32, so the MFMA operand and the lo parts use the same hi;
// lo = am x - hi rounded to f16 (the product fused into v_fma_mix), |lo| <
// 2^-10 |x| + |am - 1|, so hi + lo holds Y' to ~2^-21.  The correction operand
// (c, s) (am qk) is one fused conversion each.
__device__ __forceinline__ void y_operand(float c, float s, float am, float qk, h2v& hi,
                                          h2v& lo, h2v& q) {
  hi = __builtin_bit_cast(h2v, __builtin_amdgcn_cvt_pkrtz(c, s));
  lo[0] = (_Float16)(am * c - (float)hi[0]);
  lo[1] = (_Float16)(am * s - (float)hi[1]);
  const float aq = am * qk;
  q[0] = (_Float16)(c * aq);
  q[1] = (_Float16)(s * aq);
}

// The far-field Y operand of the 16-row kernels: as y_operand, with the
// correction operand (c, s) qk without the factor am.  That drops
// (am - 1) times the cross term, below |kappa| Qm times the cross bound
// (grid16_pairs_ok bounds the product), and one multiply per column.
__device__ __forceinline__ void y_operand_far(float c, float s, float am, float qk, h2v& hi,
                                              h2v& lo, h2v& q) {
  hi = __builtin_bit_cast(h2v, __builtin_amdgcn_cvt_pkrtz(c, s));
  lo[0] = (_Float16)__builtin_fmaf(am, c, -(float)hi[0]);
  lo[1] = (_Float16)__builtin_fmaf(am, s, -(float)hi[1]);
  q[0] = (_Float16)(c * qk);
  q[1] = (_Float16)(s * qk);
}

// MFMA operands of each tile's weights.  One wave per tile; lane l = 16 g + i
// holds row i = 2a + c, columns b = 2g, 2g+1 (re, im) of the real form of
// W / tscale, split into f16 hi and lo, and of i gx'_a W / tscale (gx' =
// 2^10 gx), with tscale[t] the power of two just above the tile's largest |w|
// component: wh = (hi, hi), wl = (lo, i gx' W / tscale) in the slot order of
// the 16x16x32 operand.
__global__ __launch_bounds__(kBlock) void k_grid_weights(
    const float2* __restrict__ wc, const float4* __restrict__ goff, uint32_t n_tiles,
    uint4* __restrict__ wh, uint4* __restrict__ wl, float* __restrict__ tscale) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * kGridWaves + (threadIdx.x >> 6);
  if (t >= n_tiles) return;  // uniform per wave
  const uint32_t i = lane & 15, g = lane >> 4, a = i >> 1, co = i & 1;
  const float2 w0 = wc[(size_t)t * 64 + 8 * a + 2 * g];
  const float2 w1 = wc[(size_t)t * 64 + 8 * a + 2 * g + 1];
  const float4 ga = goff[(size_t)t * 4 + (a >> 1)];
  const float gx = (a & 1) ? ga.y : ga.x;  // 2^10 gx_a
  float mx = fmaxf(fmaxf(fabsf(w0.x), fabsf(w0.y)), fmaxf(fabsf(w1.x), fabsf(w1.y)));
  for (int off = 32; off; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  int e = 0;
  if (mx > 0.f && __builtin_isfinite(mx)) frexpf(mx, &e);
  e = e < -126 ? -126 : (e > 127 ? 127 : e);
  const float is = ldexpf(1.f, -e);
  const float2 v0 = make_float2(w0.x * is, w0.y * is), v1 = make_float2(w1.x * is, w1.y * is);
  // real form of a complex weight V: rows (a, re) = (Vr, -Vi), (a, im) = (Vi, Vr)
  const float A[4] = {co ? v0.y : v0.x, co ? v0.x : -v0.y, co ? v1.y : v1.x,
                      co ? v1.x : -v1.y};
  // V = i gx W: Vr = -gx Wi, Vi = gx Wr
  const float C[4] = {co ? gx * v0.x : -gx * v0.y, co ? -gx * v0.y : -gx * v0.x,
                      co ? gx * v1.x : -gx * v1.y, co ? -gx * v1.y : -gx * v1.x};
  h8v hi, lo;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    _Float16 h, l;
    split_f16(A[s], h, l);
    hi[s] = h;
    hi[s + 4] = h;
    lo[s] = l;
    lo[s + 4] = (_Float16)C[s];
  }
  wh[(size_t)t * 64 + lane] = __builtin_bit_cast(uint4, hi);
  wl[(size_t)t * 64 + lane] = __builtin_bit_cast(uint4, lo);
  if (lane == 0) tscale[t] = ldexpf(1.f, e);
}

hipError_t launch_grid_weights(const float2* wc, const float4* goff, uint32_t n_tiles,
                               uint4* wh, uint4* wl, float* tscale, hipStream_t stream) {
  if (!n_tiles) return hipSuccess;
  const unsigned blocks = (n_tiles + kGridWaves - 1) / kGridWaves;
  k_grid_weights<<<blocks, kBlock, 0, stream>>>(wc, goff, n_tiles, wh, wl, tscale);
  return hipGetLastError();
}

template <int SIGN, bool OBL>
__global__ __launch_bounds__(kBlock) void k_superpose_grid(
    const double4* __restrict__ centre, const float4* __restrict__ goff,
    const int32_t* __restrict__ perm, const double* __restrict__ u,
    const float2* __restrict__ wc, const uint4* __restrict__ wh,
    const uint4* __restrict__ wl, const float* __restrict__ tscale,
    uint32_t n_tiles, uint32_t chunk, const double* __restrict__ v, uint32_t M,
    double inv_lambda, GridBounds gb, double2* __restrict__ out) {
  // per-(target, tile) constants, with the offsets scaled by 2^10 (so that
  // P' = 2^20 P, Q' = 2^20 Q): c0 = (-2^11 r_x, -2^11 r_y, 2^-20 alpha,
  // 2^-40 c2), c1 = (2^-20 kappa, f, A, 2^-20 A kappa) with A the amplitude
  // 1/d_ref (all zero when the target takes the exact path, which makes its
  // X, and so its share of the GEMM, exactly zero); separate
  // arrays keep the 16-lane reads of 16-byte rows free of bank conflicts
  __shared__ float4 s_c0[kGridWaves][64];
  __shared__ float4 s_c1[kGridWaves][64];
  __shared__ double2 s_red[kGridWaves][64][4];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t g = lane >> 4, j = lane & 15;
  const uint32_t base = blockIdx.x * kBlock + wv * 64;
  const uint32_t t0 = blockIdx.y * chunk;
  const uint32_t t1 = min(n_tiles, t0 + chunk);
  const uint32_t mo = base + lane;  // the target whose constants this lane computes
  const size_t mm = mo < M ? mo : M - 1;
  const double vx = v[3 * mm], vy = v[3 * mm + 1], vz = v[3 * mm + 2];
  double ex_r = 0.0, ex_i = 0.0;  // exact-path sum of target mo
  double ar[4], ai[4];            // target 16 B + j, rows 2g, 2g+1
  float fr[4], fi[4];
#pragma unroll
  for (int B = 0; B < 4; ++B) {
    ar[B] = ai[B] = 0.0;
    fr[B] = fi[B] = 0.f;
  }
  // Far-field check against every tile at once: if it holds for all 64
  // targets of the wave, the per-tile check is skipped.
  bool far_all;
  {
    const float rxm = (float)fmax(fabs(vx - gb.cx0), fabs(vx - gb.cx1));
    const float rym = (float)fmax(fabs(vy - gb.cy0), fabs(vy - gb.cy1));
    const double dz = fabs(vz - gb.z);
    const float idf = (float)(1.0 / dz), ia = (float)(1.0 / (dz / inv_lambda));
    const float kmax = (OBL ? 1.f : 0.5f) * idf * idf;
    const bool ok = grid_pairs_ok(2.f * rxm * gb.rho_x + gb.rho_x * gb.rho_x,
                                  2.f * rym * gb.rho_y + gb.rho_y * gb.rho_y,
                                  gb.rho_x * gb.rho_x, ia, idf, kmax);
    far_all = __all(ok);
  }
  int since = 0;
  for (uint32_t t = t0; t < t1; ++t) {
    const double4 c = centre[t];
    const float2 rho = __builtin_bit_cast(float2, c.w);  // (rho_x, rho_y)
    {
      const double rx = vx - c.x, ry = vy - c.y, rz = vz - c.z;
      double dref, id;
      hm_sqrt_inv(__builtin_fma(rz, rz, __builtin_fma(ry, ry, rx * rx)), dref, id);
      const double al = 0.5 * id * inv_lambda;
      const double id2 = id * id;
      const float c2 = (float)(-0.25 * al * id2);
      float kappa, amp;
      if constexpr (OBL) {  // |dz| / d^2 = (|r_z| / d_ref^2)(1 - q / d_ref^2 + ...)
        kappa = (float)(-id2);
        amp = (float)(__builtin_fabs(rz) * id2);
      } else {
        kappa = (float)(-0.5 * id2);
        amp = (float)id;
      }
      bool near = false;
      if (!far_all)
        near = !grid_pairs_ok(2.f * __builtin_fabsf((float)rx) * rho.x + rho.x * rho.x,
                              2.f * __builtin_fabsf((float)ry) * rho.y + rho.y * rho.y,
                              rho.x * rho.x, (float)(id * inv_lambda), (float)id, kappa);
      const float f = (float)frac_cycle(dref * inv_lambda);
      const float ks = kappa * 0x1p-20f;
      float4 c0 = make_float4((float)(-2048.0 * rx), (float)(-2048.0 * ry),
                              (float)(al * 0x1p-20), c2 * 0x1p-40f);
      float4 c1 = make_float4(ks, f, amp, amp * ks);
      if (near) {  // X = 0: the GEMM part of this target adds exactly zero
        c0 = make_float4(0.f, 0.f, 0.f, 0.f);
        c1 = c0;
      }
      s_c0[wv][lane] = c0;
      s_c1[wv][lane] = c1;
      if (near)  // rare: near field
        exact_tile_pairs<SIGN, OBL>(u, perm + (size_t)t * 64, wc + (size_t)t * 64, 64, vx, vy,
                                    vz, inv_lambda, ex_r, ex_i);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float4 go = goff[(size_t)t * 4 + g];
    const h8v Wh = __builtin_bit_cast(h8v, wh[(size_t)t * 64 + lane]);
    const h8v Wl = __builtin_bit_cast(h8v, wl[(size_t)t * 64 + lane]);
    const float ts = tscale[t];
    const float gxs[2] = {go.x, go.y}, gys[2] = {go.z, go.w};
#pragma unroll
    for (int B = 0; B < 4; ++B) {
      const float4 c0 = s_c0[wv][16 * B + j];
      const float4 c1 = s_c1[wv][16 * B + j];
      // correction operand -2 s beta r_x Q Y' 2^-10 (the weights carry
      // gx' = 2^10 gx) = s 4 pi (2^-40 c2) (-2^11 r_x) (2^20 Q) Y'
      const float kb = (float)SIGN * kFourPi * c0.w * c0.x;
      h8v y1, y2;  // (Y'_hi, Y'_lo), (Y'_hi, correction)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float gy = gys[h];
        const float Q = __builtin_fmaf(c0.y, gy, gy * gy);  // 2^20 Q
        float sn, cs;
        sincos_rev_f32(Q * __builtin_fmaf(c0.w, Q, c0.z), sn, cs);
        const float am = __builtin_fmaf(c1.x, Q, 1.f);  // 1 + kappa Q
        h2v hi, lo, yq;
        y_operand(cs, SIGN > 0 ? sn : -sn, am, Q * kb, hi, lo, yq);
        y1[2 * h] = y2[2 * h] = hi[0];
        y1[2 * h + 1] = y2[2 * h + 1] = hi[1];
        y1[4 + 2 * h] = lo[0];
        y1[5 + 2 * h] = lo[1];
        y2[4 + 2 * h] = yq[0];
        y2[5 + 2 * h] = yq[1];
      }
      f4v D = {0.f, 0.f, 0.f, 0.f};
      D = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wh, y1, D, 0, 0, 0);
      D = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wl, y2, D, 0, 0, 0);  // T / (tscale A)
      const float as = c1.z * ts, ks = c1.w * ts;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float gx = gxs[h];
        const float P = __builtin_fmaf(c0.x, gx, gx * gx);  // 2^20 P
        float sn, cs;
        sincos_rev_f32(__builtin_fmaf(P, __builtin_fmaf(c0.w, P, c0.z), c1.y), sn, cs);
        const float am = __builtin_fmaf(ks, P, as);  // tscale A (1 + kappa P)
        const float xr = am * cs, xi = SIGN > 0 ? am * sn : -(am * sn);
        fr[B] = __builtin_fmaf(xr, D[2 * h], __builtin_fmaf(-xi, D[2 * h + 1], fr[B]));
        fi[B] = __builtin_fmaf(xr, D[2 * h + 1], __builtin_fmaf(xi, D[2 * h], fi[B]));
      }
    }
    if (++since == kGridFlush) {
      since = 0;
#pragma unroll
      for (int B = 0; B < 4; ++B) {
        ar[B] += fr[B];
        ai[B] += fi[B];
        fr[B] = fi[B] = 0.f;
      }
    }
  }
#pragma unroll
  for (int B = 0; B < 4; ++B) s_red[wv][16 * B + j][g] = make_double2(ar[B] + fr[B], ai[B] + fi[B]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double sr = ex_r, si = ex_i;
#pragma unroll
  for (int q = 0; q < 4; ++q) {  // fixed order: deterministic
    const double2 p = s_red[wv][lane][q];
    sr += p.x;
    si += p.y;
  }
  if (mo < M) out[(size_t)blockIdx.y * M + mo] = make_double2(sr, si);
}

int grid_splits(size_t M, size_t n_tiles) {
  return n_tiles ? choose_splits(M, 1, n_tiles, 16) : 1;
}

hipError_t launch_superpose_grid(int sign, bool obliquity, const GridTiles& gt,
                                 const float2* wc, const uint4* wh, const uint4* wl,
                                 const float* tscale, const double* v, size_t M,
                                 double lambda, double2* y, double2* partial,
                                 size_t partial_cap, hipStream_t stream) {
  if (M == 0) return hipSuccess;
  int S = grid_splits(M, gt.n_tiles);
  while (S > 1 && (size_t)S * M > partial_cap) --S;
  const uint32_t chunk = gt.n_tiles ? (gt.n_tiles + S - 1) / S : 0;
  const dim3 grid((unsigned)((M + kBlock - 1) / kBlock), (unsigned)S);
  double2* out = S > 1 ? partial : y;
  const double il = 1.0 / lambda;
  with_sign(sign, [&](auto sg) {
    with_bool(obliquity, [&](auto obl) {
      k_superpose_grid<sg.value, (bool)obl.value><<<grid, kBlock, 0, stream>>>(
          gt.centre, gt.goff, gt.perm, gt.u, wc, wh, wl, tscale, gt.n_tiles, chunk, v,
          (uint32_t)M, il, GridBounds{gt.cx0, gt.cx1, gt.cy0, gt.cy1, gt.z, gt.rho_x, gt.rho_y},
          out);
    });
  });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_reduce_partials(S, M, partial, y, stream);
}

// ---- 16 x 16, 16 x 32 and 16 x 64 tiles on uniform lattices --------------
//
// On a lattice with constant pitches (p_x, p_y) the sources of a tile of 16
// rows (along x) and 16 NKC columns (along y, NKC = 1, 2 or 4) sit at
// gx_a = p_x k_a, gy_b = p_y k_b from its centre, k_a = a - 7.5 (a < 16),
// k_b = b - (8 NKC - 1/2) (b < 16 NKC).  A 16 x 32 tile is the pair of 16 x 16
// tiles 2T, 2T + 1 of the same rows, and a 16 x 64 tile the quad 4T .. 4T + 3,
// so all share the slots (16 a + b of each 16 x 16 tile), the weights' layout
// and the exact path's source order.
// In units of the pitch, P = p_x^2 Pt with Pt = k_a^2 + qx k_a,
// qx = -2 r_x / p_x (likewise Q, Qt, qy), and the row phase is
//   alpha P + c2 P^2 = sx k_a + ax k_a^2 + cx Pt^2,
//   sx = -2 alpha r_x p_x (cycles per pitch), ax = alpha p_x^2, cx = c2 p_x^4.
// The linear part spans several cycles over a tile, so it is evaluated
// without rounding: sx = sh + sl with sh keeping 20 significant bits (19 for
// 32 columns, 18 for 64), which makes sh k exact in FP32 (2k is an odd integer
// below 16, 32 or 64); its fraction is exact too (v_fract_f32), and only then are sl k, the
// quadratic terms and f added, so the FP32 argument of sin/cos stays below 1.5
// cycles however large the tile's phase span.  Per tile and target that leaves
// the same rounding as the 8 x 8 kernel with a quarter of the per-target
// constants and half the phasors per pair (16 x 16), 3/8 (16 x 32) or 5/16
// (16 x 64).
//
// The cross term 2 c2 P Q: its part linear in k_a goes through the GEMM (the
// correction below); of its k_a^2 part, 2 c2 p_x^2 k_a^2 Q, the share of the
// midpoint m = 28.25 of k_a^2 (0.25 .. 56.25) depends on the column only and
// joins the column phase, alpha_y = alpha + 2 c2 p_x^2 m; only
// 2 c2 p_x^2 (k_a^2 - m) Q, |k_a^2 - m| <= 28, is dropped (the guard bounds it).
//
// T for 16 targets is a 32 x 32 real product (complex 16 x 16) per 16 columns,
// computed as two row blocks rb (source rows 8 rb .. 8 rb + 7) of 16 x 16 x 32
// f16 products per chunk of 16 columns: lane l = 16 g + i holds row
// i = 2a' + c (a = 8 rb + a') of the A operand and k-slots e < 8 of column
// 16 kc + 4g + e/2, component e%2 -- the same slots as the Y operand of target
// j = l%16.  Per row block and chunk four MFMAs accumulate W_hi Y_hi,
// W_hi Y_lo, W_lo Y_hi and the correction (i k_a kCrossW W / tscale) (Y Qt kb),
// kb = s 4 pi c2 (-2 r_x) p_x p_y^2 / kCrossW; the lane receives T of rows
// 8 rb + 2g, 8 rb + 2g + 1 for target j.

constexpr int kGrid16Flush = 2;  // tiles summed in FP32 between FP64 flushes
// f16 scales of the correction operands: the weight side carries k_a kCrossW W
// (and k_a^2 kCross2W W), the column side the inverse, so that both sides stay
// in f16's normal range from a few cm to about a metre.
constexpr float kCrossW = 0x1p-4f;
constexpr float kCross2W = 0x1p-8f;
constexpr double kRowMid = 28.25;  // midpoint of k_a^2 over a tile's 16 rows
constexpr float kRowDev = 28.f;    // largest |k_a^2 - kRowMid|

// Accuracy bounds of the 16-row kernels: those of grid_pairs_ok, with the
// FP32 rounding of the phase argument bounded by its reduced size (below one
// cycle after the exact fraction, plus f and the quadratic terms).  rx2, ry2:
// largest gx^2, gy^2; rx2c: largest |gx^2 - p_x^2 kRowMid|, the dropped share
// of the cross term.  NaN fails.
__device__ __forceinline__ bool grid16_pairs_ok(float Pm, float Qm, float rx2, float ry2,
                                                float rx2c, float ia, float idf, float kappa) {
#pragma clang fp contract(off)  // the same decision in every kernel it is inlined into
  const float qmax = Pm + Qm;
  const float round_err = 6e-8f * (2.f + 0.5f * ia * (rx2 + ry2));
  const float trunc_err = 0.0625f * qmax * qmax * qmax * ia * idf * idf * idf * idf;
  const float beta = kFourPi * 0.125f * ia * idf * idf;
  const float cross = beta * Pm * Qm;
  const float ampsep = kappa * kappa * Pm * Qm;
  const float ampq = __builtin_fabsf(kappa) * Qm * cross;  // y_operand_far's dropped share
  return round_err <= kGuardCycles && trunc_err <= 0.05f * kGuardCycles && cross <= 1.5e-3f &&
         beta * rx2c * Qm <= 1e-6f && ampsep <= 1e-6f && ampq <= 1e-7f;
}

// Near-field GEMM of the 16 x 16 tiles.  The cross factor is carried to second
// order, with the q^3 terms and the quadratic amplitude: with u = p_x k_a,
// P = u^2 - 2 r_x u and
//   C = 2 c2 P Q + 3 c3 (P^2 Q + P Q^2)   (cycles),   phi = s 2 pi C,
//   e^{i phi} ~ 1 + i phi - phi^2 / 2 = 1 + i phi1 u + (i phi2 - phi1^2 / 2) u^2 + O(u^3),
//   phi1 = s 2 pi (-4 c2 r_x Q - 6 c3 r_x Q^2),  phi2 = s 2 pi (2 c2 + 12 c3 r_x^2) Q,
// the rows X_a = (1 + k1 P + k2 P^2) e^{i s 2 pi (alpha P + c2 P^2 + c3 P^3)} and
// the columns Y_b likewise in Q.  T_a = sum_b W Y + u sum_b W (i phi1 Y) +
// u^2 sum_b W (i phi2 - phi1^2 / 2) Y: the u term as the existing correction
// with both of its operands split into f16 hi and lo, the u^2 term as one more
// f16 product.  Dropped, and bounded here per (target, tile): the u^3 terms
// (12 c3 |r_x| Q u^3 in C, phi1 phi2 u^3 in phi^2 / 2), phi^3 / 6, the q^4
// phase term, the amplitude's cross term (2 k2 - k1^2) P Q and third order,
// and the f16 rounding of the u^2 operand.  NaN fails.
__device__ __forceinline__ bool grid16_near_ok(float Pm, float Qm, float arx, float rhox,
                                               float rx2, float ry2, float ia, float idf,
                                               bool obl) {
#pragma clang fp contract(off)
  const float qmax = Pm + Qm;
  const float id2 = idf * idf, id4 = id2 * id2;
  const float c2a = 0.125f * ia * id2, c3 = 0.0625f * ia * id4;  // |c2|, c3 (cycles)
  const float round_err = 6e-8f * (3.f + 0.5f * ia * (rx2 + ry2));
  const float q4 = 0.0390625f * qmax * qmax * qmax * qmax * ia * id4 * id2;
  const float r3 = rhox * rhox * rhox;
  const float u3 = 12.f * c3 * arx * Qm * r3;
  const float twopi = 0.5f * kFourPi;
  const float phi1u = twopi * (4.f * c2a + 6.f * c3 * Qm) * arx * rhox * Qm;
  const float phi2u2 = twopi * (2.f * c2a + 12.f * c3 * arx * arx) * Qm * rhox * rhox;
  const float ps = phi1u + phi2u2;
  const float ph3 = ps * ps * ps * (1.f / 6.f);
  const float mix = phi1u * phi2u2;
  const float ampx = (obl ? 1.f : 0.5f) * id4 * Pm * Qm;
  const float sq = qmax * id2;
  const float amp3 = (obl ? 1.f : 0.3125f) * sq * sq * sq;
  const float f16c2 = 1e-3f * (phi2u2 + 0.5f * phi1u * phi1u);
  return round_err <= kGuardCycles && q4 <= 0.05f * kGuardCycles &&
         u3 <= 0.05f * kGuardCycles && ph3 <= 1e-7f && mix <= 1e-7f && ampx <= 1e-7f &&
         amp3 <= 1e-7f && f16c2 <= 1e-7f;
}

// x rounded to BITS significant bits (the low 24 - BITS bits of the FP32
// significand cleared)
template <int BITS>
__device__ __forceinline__ float hi_bits(double x) {
  constexpr uint32_t mask = ~((1u << (24 - BITS)) - 1u);
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, (float)x) & mask);
}

// Near-field path of the 16-row kernels.  A (target, tile) pair that fails
// grid16_pairs_ok (targets within ~0.2 m of a 1024^2 plane: the q-series
// truncation and the first-order cross term exceed their bounds) is still far
// enough for the q-series per pair.  With r = v - c, D = |r|, the source
// (p_x k_a, p_y k_b) of the tile has q = P_a + Q_b,
//   P_a = p_x^2 k_a^2 - 2 r_x p_x k_a,  Q_b = p_y^2 k_b^2 - 2 r_y p_y k_b,
//   (d - D) / lambda = (q / (2D) - q^2 / (8 D^3) + q^3 / (16 D^5) - ...) / lambda
//   1 / d = (1 / D)(1 - q / (2 D^2) + 3 q^2 / (8 D^4) - ...)
//   |dz| / d^2 = (|r_z| / D^2)(1 - q / D^2 + q^2 / D^4 - ...)
// The linear phase s_x k_a (s_x = -2 alpha r_x p_x cycles per pitch, alpha =
// 1 / (2 D lambda), several cycles across a tile) is split as in the GEMM path:
// s_hi k_a and its fraction are exact in FP32, the rest of the phase is below
// a cycle.  Per row a: R_a = fract(sx_hi k_a) + sx_lo k_a + alpha p_x^2 k_a^2;
// per column b: C_b = fract(sy_hi k_b) + sy_lo k_b + alpha p_y^2 k_b^2 + f
// (f = frac(D / lambda)); per pair, in FP32,
//   q = P_a + Q_b,  t = R_a + C_b + q^2 (c2 + c3 q),  am = 1 + q (k1 + k2 q),
//   S_a += w_ab am e^{i s 2 pi t},
// each row's S_a added in FP64, the tile's sum scaled by A = 1 / D (or
// |r_z| / D^2).  Dropped: the q^4 phase term and the amplitude's third order.

// The series path's bounds for one (target, tile): the FP32 rounding of the
// phase (as grid16_pairs_ok), the q^4 term, 5 q^4 / (128 D^7 lambda) cycles,
// and the amplitude's third order, (q / D^2)^3 (5/16, or 1 with obliquity).
// NaN fails.
template <bool OBL>
__device__ __forceinline__ bool series_pairs_ok(float qmax, float rx2, float ry2, float ia,
                                                float idf) {
#pragma clang fp contract(off)
  const float sq = qmax * idf * idf;  // q / D^2
  const float round_err = 6e-8f * (3.f + 0.5f * ia * (rx2 + ry2));
  const float trunc4 = 0.0390625f * qmax * sq * sq * sq * ia;
  const float amp3 = (OBL ? 1.f : 0.3125f) * sq * sq * sq;
  return round_err <= kGuardCycles && trunc4 <= 0.05f * kGuardCycles && amp3 <= 1e-7f;
}

template <int BITS>
__device__ __forceinline__ float hi_bits(double x);

// The 256 NKC pairs of one tile (weights w: 256 slots per 16 x 16 part, slot
// 16 a + b) for one target, added into (ar, ai).
template <int SIGN, bool OBL, int NKC>
__device__ __forceinline__ void series_tile_pairs(const float2* __restrict__ w, double rx,
                                                  double ry, double rz, double dref, double id,
                                                  double inv_lambda, double px, double py,
                                                  double& ar, double& ai) {
#pragma clang fp contract(off)
  const double al = 0.5 * id * inv_lambda;
  const double sx = -2.0 * al * rx * px, sy = -2.0 * al * ry * py;
  const float sxh = hi_bits<20>(sx), syh = hi_bits<NKC == 1 ? 20 : (NKC == 2 ? 19 : 18)>(sy);
  const float sxl = (float)(sx - (double)sxh), syl = (float)(sy - (double)syh);
  const float ax = (float)(al * px * px), ay = (float)(al * py * py);
  const float mx = (float)(-2.0 * rx * px), my = (float)(-2.0 * ry * py);
  const float px2 = (float)(px * px), py2 = (float)(py * py);
  const float idf = (float)id, id2 = idf * idf, ia = (float)(id * inv_lambda);
  const float c2 = -0.125f * ia * id2, c3 = 0.0625f * ia * id2 * id2;
  const float k1 = OBL ? -id2 : -0.5f * id2, k2 = OBL ? id2 * id2 : 0.375f * id2 * id2;
  const float f = (float)frac_cycle(dref * inv_lambda);
  const double A = OBL ? __builtin_fabs(rz) * id * id : id;
  for (int c = 0; c < NKC; ++c) {
    float Cb[16], Qb[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const float k = (float)(16 * c + b) - (8.f * NKC - 0.5f), kk = k * k;
      Qb[b] = __builtin_fmaf(my, k, py2 * kk);
      Cb[b] = __builtin_amdgcn_fractf(syh * k) + __builtin_fmaf(syl, k, ay * kk) + f;
    }
    for (int a = 0; a < 16; ++a) {
      const float k = (float)a - 7.5f, kk = k * k;
      const float Pa = __builtin_fmaf(mx, k, px2 * kk);
      const float Ra = __builtin_amdgcn_fractf(sxh * k) + __builtin_fmaf(sxl, k, ax * kk);
      const float2* wr = w + 256 * c + 16 * a;
      float sr = 0.f, si = 0.f;
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const float2 wb = wr[b];
        const float q = Pa + Qb[b];
        const float t = __builtin_fmaf(q * q, __builtin_fmaf(c3, q, c2), Ra + Cb[b]);
        const float am = __builtin_fmaf(q, __builtin_fmaf(k2, q, k1), 1.f);
        float sn, cs;
        sincos_rev_f32(t, sn, cs);
        if (SIGN < 0) sn = -sn;
        const float pr = am * cs, pi = am * sn;
        sr = __builtin_fmaf(wb.x, pr, __builtin_fmaf(-wb.y, pi, sr));
        si = __builtin_fmaf(wb.x, pi, __builtin_fmaf(wb.y, pr, si));
      }
      ar += A * sr;
      ai += A * si;
    }
  }
}

// MFMA operands of the 16 x 16 tiles' weights, one wave per tile: for row
// block rb, wq[(6t + 3rb + q) * 64 + lane] holds, for lane l = 16 g + i,
// row i = 2a' + c (a = 8 rb + a') and columns 4g .. 4g + 3 (re, im) of the real
// form of W / tscale split into f16 hi (q = 0) and lo (q = 1), and of
// i k_a kCrossW W / tscale (q = 2, its f16 hi part).  wqn (nullable), the
// near-field GEMM's extra operands: wqn[(4t + 2rb + q) * 64 + lane] holds the
// lo part of i k_a kCrossW W / tscale (q = 0) and k_a^2 kCross2W W / tscale in
// the real form of W (q = 1).  wc: 256 slots per tile, slot 16 a + b.
// group (1, 2 or 4): tiles group T .. group T + group - 1 (the halves of a
// 16 x 32 tile, the quarters of a 16 x 64 one) share one tscale, the power of
// two above the largest of their largest components.
__global__ __launch_bounds__(kBlock) void k_grid16_weights(const float2* __restrict__ wc,
                                                           uint32_t n_tiles, uint32_t group,
                                                           uint4* __restrict__ wq,
                                                           uint4* __restrict__ wqn,
                                                           float* __restrict__ tscale) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * kGridWaves + (threadIdx.x >> 6);
  if (t >= n_tiles) return;  // uniform per wave
  const uint32_t i = lane & 15, g = lane >> 4, ap = i >> 1, co = i & 1;
  float2 w[2][4];
  float mx = 0.f;
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      w[rb][q] = wc[(size_t)t * 256 + 16 * (8 * rb + ap) + 4 * g + q];
      mx = fmaxf(mx, fmaxf(fabsf(w[rb][q].x), fabsf(w[rb][q].y)));
    }
  for (uint32_t o = 1; o < group; ++o) {  // n_tiles is a multiple of group
    const uint32_t tt = (t & ~(group - 1u)) | ((t + o) & (group - 1u));
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 x = wc[(size_t)tt * 256 + 16 * (8 * rb + ap) + 4 * g + q];
        mx = fmaxf(mx, fmaxf(fabsf(x.x), fabsf(x.y)));
      }
  }
  for (int off = 32; off; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  int e = 0;
  if (mx > 0.f && __builtin_isfinite(mx)) frexpf(mx, &e);
  e = e < -126 ? -126 : (e > 127 ? 127 : e);
  const float is = ldexpf(1.f, -e);
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const float kr = (float)(8 * rb + ap) - 7.5f;
    const float ka = kr * kCrossW, ka2 = kr * kr * kCross2W;
    h8v hi, lo, cr, crl, c2;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float vr = w[rb][q].x * is, vi = w[rb][q].y * is;
      // real form: rows (a, re) = (Vr, -Vi), (a, im) = (Vi, Vr)
      const float A0 = co ? vi : vr, A1 = co ? vr : -vi;
      // i k W: real part -k Wi, imaginary part k Wr
      const float C0 = co ? ka * vr : -ka * vi, C1 = co ? -ka * vi : -ka * vr;
      _Float16 h, l;
      split_f16(A0, h, l);
      hi[2 * q] = h;
      lo[2 * q] = l;
      split_f16(A1, h, l);
      hi[2 * q + 1] = h;
      lo[2 * q + 1] = l;
      split_f16(C0, h, l);
      cr[2 * q] = h;
      crl[2 * q] = l;
      split_f16(C1, h, l);
      cr[2 * q + 1] = h;
      crl[2 * q + 1] = l;
      // k^2 W (real scaling): the same real form as W
      c2[2 * q] = (_Float16)(ka2 * A0);
      c2[2 * q + 1] = (_Float16)(ka2 * A1);
    }
    wq[((size_t)t * 6 + 3 * rb + 0) * 64 + lane] = __builtin_bit_cast(uint4, hi);
    wq[((size_t)t * 6 + 3 * rb + 1) * 64 + lane] = __builtin_bit_cast(uint4, lo);
    wq[((size_t)t * 6 + 3 * rb + 2) * 64 + lane] = __builtin_bit_cast(uint4, cr);
    if (wqn) {
      wqn[((size_t)t * 4 + 2 * rb + 0) * 64 + lane] = __builtin_bit_cast(uint4, crl);
      wqn[((size_t)t * 4 + 2 * rb + 1) * 64 + lane] = __builtin_bit_cast(uint4, c2);
    }
  }
  if (lane == 0) tscale[t] = ldexpf(1.f, e);
}

hipError_t launch_grid16_weights(const float2* wc, uint32_t n_tiles, int group, uint4* wq,
                                 uint4* wqn, float* tscale, hipStream_t stream) {
  if (!n_tiles) return hipSuccess;
  if ((group != 1 && group != 2 && group != 4) || n_tiles % (uint32_t)group)
    return hipErrorInvalidValue;
  const unsigned blocks = (n_tiles + kGridWaves - 1) / kGridWaves;
  k_grid16_weights<<<blocks, kBlock, 0, stream>>>(wc, n_tiles, (uint32_t)group, wq, wqn,
                                                  tscale);
  return hipGetLastError();
}

// The workgroup's LDS: per-(target, tile) constants k (all zero when the
// target takes the exact path, which makes its X, and so its share of the
// GEMM, exactly zero):
//   k0 = (sx_hi, sx_lo, sy_hi, sy_lo)   linear phase, cycles per pitch
//   k1 = (ax, ay, cx, cy)               quadratic phase coefficients
//   k2 = (qx, qy, kappa p_x^2, kappa p_y^2)
//   k3 = (f, A, kb, kb3)
// and for the near-field GEMM
//   k4 = (c3 p_x^6, c3 p_y^6, k2 p_x^4, k2 p_y^4)   cubic phase, quadratic amplitude
//   k5 = (e2, 0, 0, 0)                            the k_a^2 correction's phase term
// and rr, ri: the FP64 sums of target 16 B + j, rows 8 rb + 2g + c, in slot
// [B][lane] (lane-contiguous, so free of bank conflicts; in LDS rather than
// registers: they are touched once per kGrid16Flush tiles, and the registers
// keep the kernel at 3 or 4 waves).
struct Grid16Lds {
  float4 k[6][kGridWaves][64];
  double rr[kGridWaves][4][64];
  double ri[kGridWaves][4][64];
};

// The accuracy bounds of one target against every tile at once (lateral
// extent of the tile centres, largest radii, smallest distance).
template <bool OBL>
__device__ __forceinline__ bool grid16_far(double vx, double vy, double vz, double inv_lambda,
                                           const GridBounds& gb, float rx2c) {
#pragma clang fp contract(off)
  const float rx2 = gb.rho_x * gb.rho_x, ry2 = gb.rho_y * gb.rho_y;
  const float rxm = (float)fmax(fabs(vx - gb.cx0), fabs(vx - gb.cx1));
  const float rym = (float)fmax(fabs(vy - gb.cy0), fabs(vy - gb.cy1));
  const double dz = fabs(vz - gb.z);
  const float idf = (float)(1.0 / dz), ia = (float)(1.0 / (dz / inv_lambda));
  const float kmax = (OBL ? 1.f : 0.5f) * idf * idf;
  return grid16_pairs_ok(2.f * rxm * gb.rho_x + rx2, 2.f * rym * gb.rho_y + ry2, rx2, ry2, rx2c,
                         ia, idf, kmax);
}

// Tiles t0 .. t1 - 1 of 16 x 16 NKC sources for the wave's 64 targets: centre[t],
// slots 256 NKC t .. (perm, wc), MFMA operands of chunk kc in the layout of
// 16 x 16 tile NKC t + kc (wq), scale tscale[NKC t].  far_all: every target of
// the wave passed grid16_far for these tiles (no per-tile check).  NEAR (16 x 16
// tiles only): the near-field GEMM (grid16_near_ok), with the extra operands
// wqn.  Adds into lds.rr / lds.ri (FP32 partial sums, flushed every
// kGrid16Flush tiles) and, for (target, tile) pairs that fail the bounds, into
// ex_r / ex_i.
template <int SIGN, bool OBL, int NKC, bool NEAR = false>
__device__ __forceinline__ void grid16_tiles(
    Grid16Lds& lds, const double4* __restrict__ centre, const int32_t* __restrict__ perm,
    const double* __restrict__ u, const float2* __restrict__ wc,
    const uint4* __restrict__ wq, const uint4* __restrict__ wqn,
    const float* __restrict__ tscale, uint32_t t0, uint32_t t1,
    double vx, double vy, double vz, double inv_lambda, double px, double py,
    const GridBounds& gb, const double* __restrict__ axes, const LatticeAxes& ax, bool far_all,
    double& ex_r, double& ex_i) {
  // every fused operation is written out, so that the same tiles give the same
  // bits whichever kernel this is inlined into
#pragma clang fp contract(off)
  static_assert(NKC == 1 || NKC == 2 || NKC == 4, "16, 32 or 64 columns");
  static_assert(!NEAR || NKC == 1, "the near-field GEMM takes 16 x 16 tiles");
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t g = lane >> 4, j = lane & 15;
  // this lane's rows 8 rb + 2g + c and columns 16 kc + 4g + e, as k = index - centre
  float kr[4], kc[NKC][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    kr[q] = (float)(8 * (q >> 1) + 2 * (int)g + (q & 1)) - 7.5f;
#pragma unroll
    for (int c = 0; c < NKC; ++c) kc[c][q] = (float)(16 * c + 4 * (int)g + q) - (8.f * NKC - 0.5f);
  }
  const float px2f = (float)(px * px), py2f = (float)(py * py);
  const float px4f = px2f * px2f, py4f = py2f * py2f;
  const float px6f = (float)(px * px * px * px * px * px), py6f = (float)(py * py * py * py * py * py);
  const float ipxf = (float)(1.0 / px), ipyf = (float)(1.0 / py);
  // kb = s 4 pi c2 (-2 r_x) p_x p_y^2 / kCrossW = c2 r_x kbf
  const float kbf = (float)((double)SIGN * -8.0 * M_PI * px * (py * py) / kCrossW);
  // e2 = s 2 pi (2 c2 + 12 c3 r_x^2) p_x^2 p_y^2 / kCross2W = (2 c2 + 12 c3 r_x^2) e2f
  const float e2f = (float)((double)SIGN * 2.0 * M_PI * (px * px) * (py * py) / kCross2W);
  // alpha_y = alpha (1 + mid / d^2): the row-midpoint share of the cross term
  // (the near-field GEMM carries the whole k_a^2 share in its own operand)
  const double mid = -0.5 * kRowMid * px * px;
  const float rx2 = gb.rho_x * gb.rho_x, ry2 = gb.rho_y * gb.rho_y;
  const float rx2c = kRowDev * px2f;
  float fr[4], fi[4];
#pragma unroll
  for (int B = 0; B < 4; ++B) fr[B] = fi[B] = 0.f;
  int since = 0;
  // one tile's GEMM and row phasors for the wave's 64 targets (constants in
  // lds.k), NB: with the near-field operands
  const auto tile_gemm = [&](uint32_t t, auto nb) __attribute__((always_inline)) {
    constexpr bool NB = decltype(nb)::value;
    h8v W[6 * NKC];
#pragma unroll
    for (int q = 0; q < 6 * NKC; ++q)
      W[q] = __builtin_bit_cast(h8v, wq[((size_t)t * 6 * NKC + q) * 64 + lane]);
    h8v Wn[NB ? 4 : 1];
    if constexpr (NB) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        Wn[q] = __builtin_bit_cast(h8v, wqn[((size_t)t * 4 + q) * 64 + lane]);
    }
    const float ts = tscale[(size_t)t * NKC];
#pragma unroll
    for (int B = 0; B < 4; ++B) {
      const float4 k0 = lds.k[0][wv][16 * B + j];
      const float4 k1 = lds.k[1][wv][16 * B + j];
      const float4 k2 = lds.k[2][wv][16 * B + j];
      const float4 k3 = lds.k[3][wv][16 * B + j];
      float4 k4 = make_float4(0.f, 0.f, 0.f, 0.f), k5 = k4;
      if constexpr (NB) {
        k4 = lds.k[4][wv][16 * B + j];
        k5 = lds.k[5][wv][16 * B + j];
      }
      f4v D0 = {0.f, 0.f, 0.f, 0.f}, D1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < NKC; ++c) {
        h8v yh, yl, yq, yql, y2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float k = kc[c][q], kk = k * k;
          const float Qt = __builtin_fmaf(k2.y, k, kk);
          const float r = __builtin_amdgcn_fractf(k0.z * k);  // exact
          float tq = __builtin_fmaf(k0.w, k, k1.y * kk);
          tq = __builtin_fmaf(k1.w * Qt, Qt, tq);
          if constexpr (NB) tq = __builtin_fmaf(k4.y * Qt, Qt * Qt, tq);
          float sn, cs;
          sincos_rev_f32(r + tq, sn, cs);
          if (SIGN < 0) sn = -sn;
          float am;
          if constexpr (NB)
            am = __builtin_fmaf(Qt, __builtin_fmaf(k4.w, Qt, k2.w), 1.f);  // 1 + k1 Q + k2 Q^2
          else
            am = __builtin_fmaf(k2.w, Qt, 1.f);  // 1 + kappa Q
          const float qk = NB ? Qt * __builtin_fmaf(k3.w, Qt, k3.z) : Qt * k3.z;
          h2v hi, lo, cq;  // cq: the far-field correction (the near-field GEMM forms its own)
          y_operand_far(cs, sn, am, qk, hi, lo, cq);
          yh[2 * q] = hi[0];
          yh[2 * q + 1] = hi[1];
          yl[2 * q] = lo[0];
          yl[2 * q + 1] = lo[1];
          if constexpr (NB) {
            // the u correction split into f16 hi and lo; the u^2 correction
            // (i phi2 - phi1^2 / 2) Y / kCross2W, phi1 p_x = qk kCrossW
            const float aq = am * qk;
            _Float16 h, l;
            split_f16(cs * aq, h, l);
            yq[2 * q] = h;
            yql[2 * q] = l;
            split_f16(sn * aq, h, l);
            yq[2 * q + 1] = h;
            yql[2 * q + 1] = l;
            const float re2 = -0.5f * (kCrossW * kCrossW / kCross2W) * qk * qk, im2 = k5.x * Qt;
            y2[2 * q] = (_Float16)(am * (cs * re2 - sn * im2));
            y2[2 * q + 1] = (_Float16)(am * (cs * im2 + sn * re2));
          } else {
            yq[2 * q] = cq[0];
            yq[2 * q + 1] = cq[1];
          }
        }
        const h8v* Wc = W + 6 * c;
        D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[0], yh, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[3], yh, D1, 0, 0, 0);
        D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[0], yl, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[3], yl, D1, 0, 0, 0);
        D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[1], yh, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[4], yh, D1, 0, 0, 0);
        D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[2], yq, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[5], yq, D1, 0, 0, 0);  // T / (tscale A)
        if constexpr (NB) {
          D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[2], yql, D0, 0, 0, 0);
          D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wc[5], yql, D1, 0, 0, 0);
          D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wn[0], yq, D0, 0, 0, 0);
          D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wn[2], yq, D1, 0, 0, 0);
          D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wn[1], y2, D0, 0, 0, 0);
          D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Wn[3], y2, D1, 0, 0, 0);
        }
      }
      const float at = k3.y * ts, akt = at * k2.z, ak2t = at * k4.z;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float k = kr[q], kk = k * k;
        const float Pt = __builtin_fmaf(k2.x, k, kk);
        const float r = __builtin_amdgcn_fractf(k0.x * k);  // exact
        float tp = __builtin_fmaf(k0.y, k, __builtin_fmaf(k1.x, kk, k3.x));
        tp = __builtin_fmaf(k1.z * Pt, Pt, tp);
        if constexpr (NB) tp = __builtin_fmaf(k4.x * Pt, Pt * Pt, tp);
        float sn, cs;
        sincos_rev_f32(r + tp, sn, cs);
        float am;
        if constexpr (NB)
          am = __builtin_fmaf(Pt, __builtin_fmaf(ak2t, Pt, akt), at);  // tscale A (1 + k1 P + k2 P^2)
        else
          am = __builtin_fmaf(akt, Pt, at);  // tscale A (1 + kappa P)
        const float xr = am * cs, xi = SIGN > 0 ? am * sn : -(am * sn);
        const f4v& D = (q >> 1) ? D1 : D0;
        const float tr = D[2 * (q & 1)], ti = D[2 * (q & 1) + 1];
        fr[B] = __builtin_fmaf(xr, tr, __builtin_fmaf(-xi, ti, fr[B]));
        fi[B] = __builtin_fmaf(xr, ti, __builtin_fmaf(xi, tr, fi[B]));
      }
    }
  };
  // the far-field GEMM of 16 x 64 tiles, with the column chunks outermost: one
  // chunk's weight operands live at a time (24 VGPRs; all four chunks' would
  // take 96), the four blocks' accumulators across chunks, and each block's
  // constants read from LDS where they are used (the empty asm with a memory
  // clobber keeps the compiler from holding all four blocks' constants in
  // registers).  The same operations in the same order per block as tile_gemm.
  const auto tile_gemm_far = [&](uint32_t t) __attribute__((always_inline)) {
    f4v D[4][2];
#pragma unroll
    for (int B = 0; B < 4; ++B)
#pragma unroll
      for (int h = 0; h < 2; ++h) D[B][h] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < NKC; ++c) {
      h8v W[6];
#pragma unroll
      for (int q = 0; q < 6; ++q)
        W[q] = __builtin_bit_cast(h8v, wq[((size_t)t * 6 * NKC + 6 * c + q) * 64 + lane]);
#pragma unroll
      for (int B = 0; B < 4; ++B) {
        __asm__ volatile("" ::: "memory");
        const float4 k0 = lds.k[0][wv][16 * B + j];
        const float4 k1 = lds.k[1][wv][16 * B + j];
        const float4 k2 = lds.k[2][wv][16 * B + j];
        const float4 k3 = lds.k[3][wv][16 * B + j];
        h8v yh, yl, yq;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float k = kc[c][q], kk = k * k;
          const float Qt = __builtin_fmaf(k2.y, k, kk);
          const float r = __builtin_amdgcn_fractf(k0.z * k);  // exact
          float tq = __builtin_fmaf(k0.w, k, k1.y * kk);
          tq = __builtin_fmaf(k1.w * Qt, Qt, tq);
          float sn, cs;
          sincos_rev_f32(r + tq, sn, cs);
          if (SIGN < 0) sn = -sn;
          const float am = __builtin_fmaf(k2.w, Qt, 1.f);  // 1 + kappa Q
          const float qk = Qt * k3.z;
          h2v hi, lo, cq;
          y_operand_far(cs, sn, am, qk, hi, lo, cq);
          yh[2 * q] = hi[0];
          yh[2 * q + 1] = hi[1];
          yl[2 * q] = lo[0];
          yl[2 * q + 1] = lo[1];
          yq[2 * q] = cq[0];
          yq[2 * q + 1] = cq[1];
        }
        D[B][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[0], yh, D[B][0], 0, 0, 0);
        D[B][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[3], yh, D[B][1], 0, 0, 0);
        D[B][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[0], yl, D[B][0], 0, 0, 0);
        D[B][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[3], yl, D[B][1], 0, 0, 0);
        D[B][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[1], yh, D[B][0], 0, 0, 0);
        D[B][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[4], yh, D[B][1], 0, 0, 0);
        D[B][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[2], yq, D[B][0], 0, 0, 0);
        D[B][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(W[5], yq, D[B][1], 0, 0, 0);
      }
    }
    const float ts = tscale[(size_t)t * NKC];
#pragma unroll
    for (int B = 0; B < 4; ++B) {
      __asm__ volatile("" ::: "memory");
      const float4 k0 = lds.k[0][wv][16 * B + j];
      const float4 k1 = lds.k[1][wv][16 * B + j];
      const float4 k2 = lds.k[2][wv][16 * B + j];
      const float4 k3 = lds.k[3][wv][16 * B + j];
      const float at = k3.y * ts, akt = at * k2.z;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float k = kr[q], kk = k * k;
        const float Pt = __builtin_fmaf(k2.x, k, kk);
        const float r = __builtin_amdgcn_fractf(k0.x * k);  // exact
        float tp = __builtin_fmaf(k0.y, k, __builtin_fmaf(k1.x, kk, k3.x));
        tp = __builtin_fmaf(k1.z * Pt, Pt, tp);
        float sn, cs;
        sincos_rev_f32(r + tp, sn, cs);
        const float am = __builtin_fmaf(akt, Pt, at);  // tscale A (1 + kappa P)
        const float xr = am * cs, xi = SIGN > 0 ? am * sn : -(am * sn);
        const f4v& Dq = (q >> 1) ? D[B][1] : D[B][0];
        const float tr = Dq[2 * (q & 1)], ti = Dq[2 * (q & 1) + 1];
        fr[B] = __builtin_fmaf(xr, tr, __builtin_fmaf(-xi, ti, fr[B]));
        fi[B] = __builtin_fmaf(xr, ti, __builtin_fmaf(xi, tr, fi[B]));
      }
    }
  };
  for (uint32_t t = t0; t < t1; ++t) {
    // near_mode (wave-uniform, NEAR only): some target of the wave fails the
    // far-field bounds of this tile, so the wave takes the near-field GEMM for it
    bool near_mode = false;
    {
      const double4 c = centre[t];
      const double rx = vx - c.x, ry = vy - c.y, rz = vz - c.z;
      double dref, id;
      hm_sqrt_inv(__builtin_fma(rz, rz, __builtin_fma(ry, ry, rx * rx)), dref, id);
      const double al = 0.5 * id * inv_lambda;
      const float idf = (float)id, alf = (float)al;
      const float rxf = (float)rx, ryf = (float)ry;
      const float id2f = idf * idf;
      const float c2f = -0.25f * alf * id2f;
      const float kf = OBL ? -id2f : -0.5f * id2f;
      const float amp = OBL ? __builtin_fabsf((float)rz) * id2f : idf;
      const float Pm = 2.f * __builtin_fabsf(rxf) * gb.rho_x + rx2,
                  Qm = 2.f * __builtin_fabsf(ryf) * gb.rho_y + ry2;
      const bool far_ok = far_all || grid16_pairs_ok(Pm, Qm, rx2, ry2, rx2c, 2.f * alf, idf, kf);
      bool near = !far_ok;
      if constexpr (NEAR) {
        near_mode = __builtin_amdgcn_ballot_w64(!far_ok) != 0;
        if (near_mode)
          near = !grid16_near_ok(Pm, Qm, __builtin_fabsf(rxf), gb.rho_x, rx2, ry2, 2.f * alf,
                                 idf, OBL);
      }
      // the linear phase needs FP64 (it spans several cycles and is split
      // exactly); the small coefficients are FP32 products
      const double aly = al * __builtin_fma(near_mode ? 0.0 : mid, id * id, 1.0);
      const double sx = -2.0 * al * rx * px, sy = -2.0 * aly * ry * py;
      const float sxh = hi_bits<20>(sx), syh = hi_bits<NKC == 1 ? 20 : (NKC == 2 ? 19 : 18)>(sy);
      float4 k0 = make_float4(sxh, (float)(sx - (double)sxh), syh, (float)(sy - (double)syh));
      float4 k1 = make_float4(alf * px2f, (float)aly * py2f, c2f * px4f, c2f * py4f);
      float4 k2 = make_float4(-2.f * rxf * ipxf, -2.f * ryf * ipyf, kf * px2f, kf * py2f);
      const float kb = c2f * rxf * kbf;
      float4 k3 = make_float4((float)frac_cycle(dref * inv_lambda), amp, kb, 0.f);
      float4 k4 = make_float4(0.f, 0.f, 0.f, 0.f), k5 = k4;
      if (NEAR && near_mode) {
        // c3 = 1 / (16 D^5 lambda), k2 = 3 / (8 D^4) (|dz| / d^2: 1 / D^4)
        const float c3f = 0.0625f * (2.f * alf) * id2f * id2f;
        const float k2a = (OBL ? 1.f : 0.375f) * id2f * id2f;
        k4 = make_float4(c3f * px6f, c3f * py6f, k2a * px4f, k2a * py4f);
        // the q^2 part of the u correction: 6 c3 / (4 c2) p_y^2 = -0.75 p_y^2 / D^2
        k3.w = -0.75f * kb * id2f * py2f;
        k5.x = (2.f * c2f + 12.f * c3f * rxf * rxf) * e2f;
      }
      if (near) {
        k0 = make_float4(0.f, 0.f, 0.f, 0.f);
        k1 = k2 = k3 = k4 = k5 = k0;
      }
      lds.k[0][wv][lane] = k0;
      lds.k[1][wv][lane] = k1;
      lds.k[2][wv][lane] = k2;
      lds.k[3][wv][lane] = k3;
      if (NEAR && near_mode) {
        lds.k[4][wv][lane] = k4;
        lds.k[5][wv][lane] = k5;
      }
      if (near) {
        if (gb.series && series_pairs_ok<OBL>(Pm + Qm, rx2, ry2, 2.f * alf, idf))
          series_tile_pairs<SIGN, OBL, NKC>(wc + (size_t)t * 256 * NKC, rx, ry, rz, dref, id,
                                            inv_lambda, px, py, ex_r, ex_i);
        else if (NKC == 1 && axes) {
          // tile t = I tj + J holds rows 16 I .. and columns 16 J .. of the lattice
          // (wave-uniform, kept in scalar registers so that xs and ys take scalar loads)
          const uint32_t I = __builtin_amdgcn_readfirstlane(t / ax.tj), J = t - I * ax.tj;
          const int ni = min(16, (int)ax.nx - 16 * (int)I);
          const int nj = min(16, (int)ax.ny - 16 * (int)J);
          if (nj > 0)
            exact_lattice_tile<SIGN, OBL>(axes + 16 * (size_t)I, axes + ax.ys + 16 * (size_t)J,
                                          gb.z, ni, nj, wc + (size_t)t * 256, vx, vy, vz,
                                          inv_lambda, ex_r, ex_i);
        } else
          exact_tile_pairs<SIGN, OBL>(u, perm + (size_t)t * 256 * NKC,
                                      wc + (size_t)t * 256 * NKC, 256 * NKC, vx, vy, vz,
                                      inv_lambda, ex_r, ex_i);
      }
      // a wave whose targets are all near adds nothing through the GEMM
      if (__builtin_amdgcn_ballot_w64(!near) == 0) continue;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if constexpr (NEAR) {
      if (near_mode)
        tile_gemm(t, std::true_type{});
      else
        tile_gemm(t, std::false_type{});
    } else if constexpr (NKC == 4) {
      tile_gemm_far(t);
    } else {
      tile_gemm(t, std::false_type{});
    }
    if (++since == kGrid16Flush) {
      since = 0;
#pragma unroll
      for (int B = 0; B < 4; ++B) {
        lds.rr[wv][B][lane] += fr[B];
        lds.ri[wv][B][lane] += fi[B];
        fr[B] = fi[B] = 0.f;
      }
    }
  }
#pragma unroll
  for (int B = 0; B < 4; ++B) {
    lds.rr[wv][B][lane] += fr[B];
    lds.ri[wv][B][lane] += fi[B];
  }
}

// One wave per 64 targets.  NKC = 1: the 16 x 16 tiles t0 .. t1 - 1 of centre
// and gb.  NKC = 2: the 16 x 32 tiles of centre and gb when every target of
// the wave passes their far-field bounds; otherwise the wave takes the same
// sources as the 16 x 16 tiles 2 t0 .. 2 t1 - 1 (centre16, gb16), whose
// looser bounds send fewer (target, tile) pairs to the exact path.  NKC = 4:
// the 16 x 64 tiles of centre and gb likewise, else the 16 x 32 tiles
// 2 t0 .. 2 t1 - 1 (centre32, gb32) when the wave passes their bounds, else
// the 16 x 16 tiles 4 t0 .. 4 t1 - 1.  The n_tail sources after a padded
// lattice (tail_perm, tail_wc) are summed exactly by every target in the
// first source split.
template <int SIGN, bool OBL, int NKC>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_superpose_grid16(
    const double4* __restrict__ centre, const int32_t* __restrict__ perm,
    const double* __restrict__ u, const float2* __restrict__ wc,
    const uint4* __restrict__ wq, const uint4* __restrict__ wqn,
    const float* __restrict__ tscale, uint32_t n_tiles,
    uint32_t chunk, const double* __restrict__ v, uint32_t M, double inv_lambda, double px,
    double py, GridBounds gb, const double4* __restrict__ centre32, GridBounds gb32,
    const double4* __restrict__ centre16, GridBounds gb16, const double* __restrict__ axes,
    LatticeAxes ax,
    const int32_t* __restrict__ tail_perm, const float2* __restrict__ tail_wc, uint32_t n_tail,
    double2* __restrict__ out) {
  __shared__ Grid16Lds lds;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t base = blockIdx.x * kBlock + wv * 64;
  const uint32_t t0 = blockIdx.y * chunk;
  const uint32_t t1 = min(n_tiles, t0 + chunk);
  const uint32_t mo = base + lane;
  const size_t mm = mo < M ? mo : M - 1;
  const double vx = v[3 * mm], vy = v[3 * mm + 1], vz = v[3 * mm + 2];
#pragma unroll
  for (int B = 0; B < 4; ++B) {
    lds.rr[wv][B][lane] = 0.0;
    lds.ri[wv][B][lane] = 0.0;
  }
  const float rx2c = kRowDev * (float)(px * px);
  double ex_r = 0.0, ex_i = 0.0;
  const bool far = __all(grid16_far<OBL>(vx, vy, vz, inv_lambda, gb, rx2c));
  bool far32 = false;
  if constexpr (NKC == 4)
    far32 = !far && __all(grid16_far<OBL>(vx, vy, vz, inv_lambda, gb32, rx2c));
  if (far) {
    grid16_tiles<SIGN, OBL, NKC>(lds, centre, perm, u, wc, wq, wqn, tscale, t0, t1, vx, vy, vz,
                                 inv_lambda, px, py, gb, axes, ax, true, ex_r, ex_i);
  } else if (far32) {
    grid16_tiles<SIGN, OBL, 2>(lds, centre32, perm, u, wc, wq, wqn, tscale, 2 * t0, 2 * t1, vx,
                               vy, vz, inv_lambda, px, py, gb32, axes, ax, true, ex_r, ex_i);
  } else {
    // 16 x 16 tiles: those of the launch (NKC = 1), or the halves of its 16 x 32
    // tiles, whose looser bounds keep more pairs on the far-field GEMM
    const double4* c16 = NKC == 1 ? centre : centre16;
    const GridBounds& b16 = NKC == 1 ? gb : gb16;
    const uint32_t s0 = NKC * t0, s1 = NKC * t1;
    const bool far16 = NKC == 1 ? false : __all(grid16_far<OBL>(vx, vy, vz, inv_lambda, gb16, rx2c));
    if (far16 || !b16.near_gemm)
      grid16_tiles<SIGN, OBL, 1>(lds, c16, perm, u, wc, wq, wqn, tscale, s0, s1, vx, vy, vz,
                                 inv_lambda, px, py, b16, axes, ax, far16, ex_r, ex_i);
    else  // the near-field GEMM, with the series and exact paths beyond its bounds
      grid16_tiles<SIGN, OBL, 1, true>(lds, c16, perm, u, wc, wq, wqn, tscale, s0, s1, vx, vy,
                                       vz, inv_lambda, px, py, b16, axes, ax, false, ex_r, ex_i);
  }
  // in blocks of 256 sources, so that each FP32 partial sum stays short
  if (blockIdx.y == 0)
    for (uint32_t s0 = 0; s0 < n_tail; s0 += 256)
      exact_tile_pairs<SIGN, OBL>(u, tail_perm + s0, tail_wc + s0, (int)min(256u, n_tail - s0),
                                  vx, vy, vz, inv_lambda, ex_r, ex_i);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double sr = ex_r, si = ex_i;
#pragma unroll
  for (int q = 0; q < 4; ++q) {  // fixed order: deterministic
    sr += lds.rr[wv][lane >> 4][16 * q + (lane & 15)];
    si += lds.ri[wv][lane >> 4][16 * q + (lane & 15)];
  }
  if (mo < M) out[(size_t)blockIdx.y * M + mo] = make_double2(sr, si);
}

hipError_t launch_superpose_grid16(int sign, bool obliquity, int nkc, const GridTiles& gt,
                                   const GridTiles& pairs, const GridTiles& halves,
                                   const float2* wc, const uint4* wq, const uint4* wqn,
                                   const float* tscale, const double* v, size_t M,
                                   double lambda, double2* y, double2* partial,
                                   size_t partial_cap, hipStream_t stream) {
  if (M == 0) return hipSuccess;
  if (nkc != 1 && nkc != 2 && nkc != 4) return hipErrorInvalidValue;
  if (nkc > 1 && (halves.n_tiles != (uint32_t)nkc * gt.n_tiles || !halves.centre))
    return hipErrorInvalidValue;
  if (nkc == 4 && (pairs.n_tiles != 2 * gt.n_tiles || !pairs.centre))
    return hipErrorInvalidValue;
  if (gt.n_tail && (!gt.tail_perm || !gt.tail_wc)) return hipErrorInvalidValue;
  int S = grid_splits(M, gt.n_tiles);
  while (S > 1 && (size_t)S * M > partial_cap) --S;
  const uint32_t chunk = gt.n_tiles ? (gt.n_tiles + S - 1) / S : 0;
  const dim3 grid((unsigned)((M + kBlock - 1) / kBlock), (unsigned)S);
  double2* out = S > 1 ? partial : y;
  const double il = 1.0 / lambda;
  const int series = tuning().grid_series ? 1 : 0;
  const int near_gemm = tuning().grid_near && wqn ? 1 : 0;
  const GridBounds gb{gt.cx0,  gt.cx1,   gt.cy0,   gt.cy1, gt.z,
                      gt.rho_x, gt.rho_y, series, near_gemm};
  const GridBounds gb16{halves.cx0,   halves.cx1,   halves.cy0, halves.cy1, halves.z,
                        halves.rho_x, halves.rho_y, series,     near_gemm};
  const GridBounds gb32{pairs.cx0,   pairs.cx1,   pairs.cy0, pairs.cy1, pairs.z,
                        pairs.rho_x, pairs.rho_y, series,    near_gemm};
  // the exact path's row and column tables (16 x 16 tiles)
  const double* axes = halves.xs;
  const LatticeAxes ax{(uint32_t)(halves.ys - halves.xs), halves.tj, halves.nx, halves.ny};
  if (axes && (!halves.ys || !ax.tj ||
               (uint64_t)ax.tj * ((ax.nx + 15) / 16) !=
                   (uint64_t)(nkc == 1 ? gt.n_tiles : halves.n_tiles) ||
               ax.ys != 16 * ((ax.nx + 15) / 16)))
    return hipErrorInvalidValue;
  with_sign(sign, [&](auto sg) {
    with_bool(obliquity, [&](auto obl) {
      if (nkc == 4)
        k_superpose_grid16<sg.value, (bool)obl.value, 4><<<grid, kBlock, 0, stream>>>(
            gt.centre, gt.perm, gt.u, wc, wq, wqn, tscale, gt.n_tiles, chunk, v, (uint32_t)M,
            il, gt.px, gt.py, gb, pairs.centre, gb32, halves.centre, gb16, axes, ax,
            gt.tail_perm, gt.tail_wc, gt.n_tail, out);
      else if (nkc == 2)
        k_superpose_grid16<sg.value, (bool)obl.value, 2><<<grid, kBlock, 0, stream>>>(
            gt.centre, gt.perm, gt.u, wc, wq, wqn, tscale, gt.n_tiles, chunk, v, (uint32_t)M,
            il, gt.px, gt.py, gb, gt.centre, gb, halves.centre, gb16, axes, ax, gt.tail_perm,
            gt.tail_wc, gt.n_tail, out);
      else
        k_superpose_grid16<sg.value, (bool)obl.value, 1><<<grid, kBlock, 0, stream>>>(
            gt.centre, gt.perm, gt.u, wc, wq, wqn, tscale, gt.n_tiles, chunk, v, (uint32_t)M,
            il, gt.px, gt.py, gb, gt.centre, gb, gt.centre, gb, axes, ax, gt.tail_perm,
            gt.tail_wc, gt.n_tail, out);
    });
  });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_reduce_partials(S, M, partial, y, stream);
}

}  // namespace holo
