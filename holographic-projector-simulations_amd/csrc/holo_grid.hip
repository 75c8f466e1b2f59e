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
// and T for 16 targets at once is two 16 x 16 real products (complex 8 x 8
// written out) into one accumulator: the main term on v_mfma_f32_16x16x4f32
// in FP32 with the tile's weights divided by a power of two tscale, the
// correction -- below 1e-4 of it -- on v_mfma_f32_16x16x16f16 accumulating
// onto the FP32 result; X carries the factor tscale back.
//
// Work split (wave of 64 lanes, 64 targets): lane l = 16 g + j computes the
// FP64 constants of target l for the tile and shares them through LDS; for
// each block B of 16 targets it then supplies Y of columns 2g, 2g+1 of target
// 16 B + j as the GEMM's B operand and receives Z of rows 2g, 2g+1, which it
// multiplies by its X.  Operand layout (16x16x4: lane l holds A[l%16][l/16],
// B[l/16][l%16], D[4(l/16)+r][l%16]; 16x16x16 holds four consecutive k):
//   k = 4g + s  <->  column b = 2g + s/2, component s%2 (re, im)
//   i = 2a + c  <->  row a, component c
// so the weights enter as A[2a+c][2b+c'] = (Wr, -Wi; Wi, Wr) (k_grid_weights).
#include <hip/hip_runtime.h>
#include <math.h>

#include "holo_device.hpp"
#include "holo_internal.hpp"

namespace holo {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h4v __attribute__((ext_vector_type(4)));

constexpr int kGridWaves = kBlock / 64;
constexpr int kGridFlush = 8;  // tiles summed in FP32 between FP64 flushes
constexpr float kFourPi = 12.566370614359172f;

// Lattice-wide bounds for the once-per-launch accuracy check.
struct GridBounds {
  double cx0, cx1, cy0, cy1, z;
  float rho_x, rho_y;
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

// MFMA operands of each tile's weights.  One wave per tile; lane l = 16 g + i
// holds row i = 2a + c, columns b = 2g, 2g+1 (re, im) of the real form of
// W / tscale (FP32) and of i gx'_a W / tscale (f16, gx' = 2^10 gx), with
// tscale[t] the power of two just above the tile's largest |w| component.
__global__ __launch_bounds__(kBlock) void k_grid_weights(
    const float2* __restrict__ wc, const float4* __restrict__ goff, uint32_t n_tiles,
    float4* __restrict__ wa, uint2* __restrict__ wa16, float* __restrict__ tscale) {
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
  float4 A;
  A.x = co ? v0.y : v0.x;
  A.y = co ? v0.x : -v0.y;
  A.z = co ? v1.y : v1.x;
  A.w = co ? v1.x : -v1.y;
  // V = i gx W: Vr = -gx Wi, Vi = gx Wr
  const float4 C = make_float4(co ? gx * v0.x : -gx * v0.y, co ? -gx * v0.y : -gx * v0.x,
                               co ? gx * v1.x : -gx * v1.y, co ? -gx * v1.y : -gx * v1.x);
  const h4v h = {(_Float16)C.x, (_Float16)C.y, (_Float16)C.z, (_Float16)C.w};
  wa[(size_t)t * 64 + lane] = A;
  wa16[(size_t)t * 64 + lane] = __builtin_bit_cast(uint2, h);
  if (lane == 0) tscale[t] = ldexpf(1.f, e);
}

hipError_t launch_grid_weights(const float2* wc, const float4* goff, uint32_t n_tiles,
                               float4* wa, uint2* wa16, float* tscale, hipStream_t stream) {
  if (!n_tiles) return hipSuccess;
  const unsigned blocks = (n_tiles + kGridWaves - 1) / kGridWaves;
  k_grid_weights<<<blocks, kBlock, 0, stream>>>(wc, goff, n_tiles, wa, wa16, tscale);
  return hipGetLastError();
}

template <int SIGN, bool OBL>
__global__ __launch_bounds__(kBlock) void k_superpose_grid(
    const double4* __restrict__ centre, const float4* __restrict__ goff,
    const int32_t* __restrict__ perm, const double* __restrict__ u,
    const float2* __restrict__ wc, const float4* __restrict__ wa,
    const uint2* __restrict__ wa16, const float* __restrict__ tscale,
    uint32_t n_tiles, uint32_t chunk, const double* __restrict__ v, uint32_t M,
    double inv_lambda, GridBounds gb, double2* __restrict__ out) {
  // per-(target, tile) constants, with the offsets scaled by 2^10 (so that
  // P' = 2^20 P, Q' = 2^20 Q): c0 = (-2^11 r_x, -2^11 r_y, 2^-20 alpha,
  // 2^-40 c2), c1 = (2^-20 kappa, f, A, 2^-20 A kappa) with A the amplitude
  // 1/d_ref (all zero when the target takes the exact path); separate
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
      hm_sqrt_inv_fast(__builtin_fma(rz, rz, __builtin_fma(ry, ry, rx * rx)), dref, id);
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
      if (near) {  // the GEMM part of this target gets exactly zero
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
    const float4 A = wa[(size_t)t * 64 + lane];
    const h4v A16 = __builtin_bit_cast(h4v, wa16[(size_t)t * 64 + lane]);
    const float ts = tscale[t];
    const float gxs[2] = {go.x, go.y}, gys[2] = {go.z, go.w};
#pragma unroll
    for (int B = 0; B < 4; ++B) {
      const float4 c0 = s_c0[wv][16 * B + j];
      const float4 c1 = s_c1[wv][16 * B + j];
      // f16 operand -2 s beta r_x Q Y 2^-10 (the weights carry gx' = 2^10 gx)
      // = s 4 pi (2^-40 c2) (-2^11 r_x) (2^20 Q) Y
      const float kb = (float)SIGN * kFourPi * c0.w * c0.x;
      float yv[4];
      h4v y16;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float gy = gys[h];
        const float Q = __builtin_fmaf(c0.y, gy, gy * gy);  // 2^20 Q
        float sn, cs;
        sincos_rev_f32(Q * __builtin_fmaf(c0.w, Q, c0.z), sn, cs);
        const float am = __builtin_fmaf(c1.w, Q, c1.z);  // A (1 + kappa Q)
        const float yr = am * cs, yi = SIGN > 0 ? am * sn : -(am * sn);
        yv[2 * h] = yr;
        yv[2 * h + 1] = yi;
        const float qk = Q * kb;
        y16[2 * h] = (_Float16)(yr * qk);
        y16[2 * h + 1] = (_Float16)(yi * qk);
      }
      f4v D = {0.f, 0.f, 0.f, 0.f};
      D = __builtin_amdgcn_mfma_f32_16x16x4f32(A.x, yv[0], D, 0, 0, 0);
      D = __builtin_amdgcn_mfma_f32_16x16x4f32(A.y, yv[1], D, 0, 0, 0);
      D = __builtin_amdgcn_mfma_f32_16x16x4f32(A.z, yv[2], D, 0, 0, 0);
      D = __builtin_amdgcn_mfma_f32_16x16x4f32(A.w, yv[3], D, 0, 0, 0);
      D = __builtin_amdgcn_mfma_f32_16x16x16f16(A16, y16, D, 0, 0, 0);  // T / tscale
      const float ks = c1.x * ts;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float gx = gxs[h];
        const float P = __builtin_fmaf(c0.x, gx, gx * gx);  // 2^20 P
        float sn, cs;
        sincos_rev_f32(__builtin_fmaf(P, __builtin_fmaf(c0.w, P, c0.z), c1.y), sn, cs);
        const float am = __builtin_fmaf(ks, P, ts);  // tscale (1 + kappa P)
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
                                 const float2* wc, const float4* wa, const uint2* wa16,
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
          gt.centre, gt.goff, gt.perm, gt.u, wc, wa, wa16, tscale, gt.n_tiles, chunk, v,
          (uint32_t)M, il, GridBounds{gt.cx0, gt.cx1, gt.cy0, gt.cy1, gt.z, gt.rho_x, gt.rho_y},
          out);
    });
  });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_reduce_partials(S, M, partial, y, stream);
}

}  // namespace holo
