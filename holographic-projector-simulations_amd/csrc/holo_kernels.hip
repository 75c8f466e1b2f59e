// Hand-written gfx950 kernels for the Huygens point-source superposition
//   y_m = sum_n a_n / d_nm * exp(i (phi_n +/- 2 pi d_nm / lambda))
// (cuda/superposition.cu:35-154 computes the same sum with per_block and a
// CUB/cuBLAS reduction; these kernels are a new design for CDNA4):
//
// * One lane owns R targets; their positions and FP64 accumulators stay in
//   VGPRs for the whole launch.
// * The source index is wave-uniform, so source data is fetched with scalar
//   loads into SGPRs and broadcast to all 64 lanes for free -- no LDS and no
//   per-lane source traffic.
// * Sources are split into S contiguous chunks (blockIdx.y) only when the
//   target count alone cannot fill the 256 CUs; the S partial fields are
//   summed afterwards in a fixed order, so results are deterministic.
// * No MFMA: the sum is not a contraction; it is bound by VALU and
//   transcendental issue.
#include <hip/hip_runtime.h>
#include <math.h>

#include "holo_device.hpp"
#include "holo_internal.hpp"

namespace holo {

Tuning& tuning() {
  static Tuning t;
  return t;
}

int choose_splits(size_t M, int R, size_t units, size_t min_units, size_t max_splits) {
  const Tuning& tu = tuning();
  if (tu.splits > 0) return (int)std::min<size_t>(tu.splits, std::max<size_t>(units, 1));
  const size_t waves = (M + 64 * R - 1) / (64 * R);
  size_t S = (tu.target_waves + waves - 1) / std::max<size_t>(waves, 1);
  const size_t max_by_units = std::max<size_t>(units / std::max<size_t>(min_units, 1), 1);
  S = std::min(S, max_by_units);
  S = std::min<size_t>(S, max_splits);
  return (int)std::max<size_t>(S, 1);
}

// ---------------------------------------------------------------------------
// FP64 brute force.

// PZ: every source has the z of source 0, so dz = vz - uz is formed once per
// target (the same bits as per pair).
// ROWS (with PZ): the sources come in rows of `row` consecutive sources that
// share one x (a lattice in x-slow order, cuda/init.h:89-158), so dx^2 is
// formed once per (target, row): again the same bits as per pair.
// s_tab: the sin/cos table, already in LDS.
template <int SIGN, int R, int WF, int MATH, bool OBL, bool PZ, bool ROWS>
__device__ __forceinline__ void superpose_f64_body(
    const double* __restrict__ w0, const double* __restrict__ w1,
    const double* __restrict__ u, uint32_t N, uint32_t chunk, uint32_t row,
    const double* __restrict__ v, uint32_t M, double K,
    double2* __restrict__ out, const double* s_tab) {
  static_assert(!ROWS || PZ, "rows need sources in one plane");
  const uint32_t n0 = blockIdx.y * chunk;
  const uint32_t n1 = min(N, n0 + chunk);
  const uint32_t base = blockIdx.x * (kBlock * R) + threadIdx.x;
  double vx[R], vy[R], vz[R], ar[R], ai[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t m = base + r * kBlock;
    const size_t mm = m < M ? m : M - 1;
    vx[r] = v[3 * mm];
    vy[r] = v[3 * mm + 1];
    vz[r] = v[3 * mm + 2];
    ar[r] = 0.0;
    ai[r] = 0.0;
  }
  double dz0[R];
#pragma unroll
  for (int r = 0; r < R; ++r) dz0[r] = PZ && N ? vz[r] - u[2] : 0.0;
  // sources [a, b); with ROWS their dx^2 per target is dxx
  auto sources = [&](uint32_t a, uint32_t b, const double(&dxx)[R]) {
    for (uint32_t n = a; n < b; ++n) {
      const double ux = ROWS ? 0.0 : u[3 * (size_t)n];
      const double uy = u[3 * (size_t)n + 1];
      const double uz = PZ ? 0.0 : u[3 * (size_t)n + 2];
      if constexpr (WF == kPolar) {
        const double a = w0[n];
        const double phi = w1[n];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double dy = vy[r] - uy, dz = PZ ? dz0[r] : vz[r] - uz;
          PairF64 p;
          if constexpr (ROWS)
            p = pair_f64_s<SIGN, MATH, OBL>(a, phi, dxx[r], dy, dz, K);
          else
            p = pair_f64_d<SIGN, MATH, OBL>(a, phi, vx[r] - ux, dy, dz, K);
          accumulate_phasor<MATH>(p.r, p.arg, ar[r], ai[r], s_tab);
        }
      } else {
        const double wr = w0[2 * (size_t)n];
        const double wi = w0[2 * (size_t)n + 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double dx = vx[r] - ux, dy = vy[r] - uy, dz = PZ ? dz0[r] : vz[r] - uz;
          double d, id;
          hm_sqrt_inv(__builtin_fma(dz, dz, __builtin_fma(dy, dy, ROWS ? dxx[r] : dx * dx)),
                      d, id);
          if constexpr (OBL) id *= __builtin_fabs(dz) * id;
          const double arg = (SIGN > 0 ? d : -d) * K;
          double s, c;
          if constexpr (MATH == 0) {
            hm_sincos(arg, s_tab, s, c);
          } else {
            sincos(arg, &s, &c);
          }
          c *= id;
          s *= id;
          ar[r] += wr * c - wi * s;
          ai[r] += wr * s + wi * c;
        }
      }
    }
  };
  if constexpr (ROWS) {
    for (uint32_t n = n0; n < n1;) {
      const uint32_t e = min(n1, (n / row + 1) * row);  // end of n's row
      const double ux = u[3 * (size_t)n];
      double dxx[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double dx = vx[r] - ux;
        dxx[r] = dx * dx;
      }
      sources(n, e, dxx);
      n = e;
    }
  } else {
    const double none[R] = {};
    sources(n0, n1, none);
  }
  double2* dst = out + (size_t)blockIdx.y * M;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t m = base + r * kBlock;
    if (m < M) dst[m] = make_double2(ar[r], ai[r]);
  }
}

// The sin/cos table, 32 KB, into LDS (5 workgroups per CU, as the VGPRs allow).
constexpr int kTab = sizeof(kSinCosTab) / sizeof(double);
static_assert(kTab % kBlock == 0, "table copy");

__device__ __forceinline__ void load_sincos_table(double* s_tab) {
#pragma unroll
  for (int q = 0; q < kTab / kBlock; ++q)
    s_tab[threadIdx.x + q * kBlock] = kSinCosTab[threadIdx.x + q * kBlock];
  __syncthreads();
}

template <int SIGN, int R, int WF, int MATH, bool OBL, bool PZ, bool ROWS>
__global__ __launch_bounds__(kBlock) void k_superpose_f64(
    const double* __restrict__ w0, const double* __restrict__ w1,
    const double* __restrict__ u, uint32_t N, uint32_t chunk, uint32_t row,
    const double* __restrict__ v, uint32_t M, double K,
    double2* __restrict__ out) {
  __shared__ double s_tab[kTab];
  load_sincos_table(s_tab);
  superpose_f64_body<SIGN, R, WF, MATH, OBL, PZ, ROWS>(w0, w1, u, N, chunk, row, v, M, K, out,
                                                        s_tab);
}

// The same sum with the source layout read from device memory (written by
// launch_source_layout): layout[0] bit 0 = every source has the z of source
// 0, bit 1 = rows of layout[1] sources with one x.  The branch is uniform over
// the launch, and every branch gives the same bits.
template <int SIGN, int R, int WF, bool OBL>
__global__ __launch_bounds__(kBlock) void k_superpose_f64_auto(
    const double* __restrict__ w0, const double* __restrict__ w1,
    const double* __restrict__ u, uint32_t N, uint32_t chunk,
    const uint32_t* __restrict__ layout, const double* __restrict__ v, uint32_t M, double K,
    double2* __restrict__ out) {
  __shared__ double s_tab[kTab];
  load_sincos_table(s_tab);
  const uint32_t mode = layout[0], row = layout[1];
  if ((mode & 3u) == 3u)
    superpose_f64_body<SIGN, R, WF, 0, OBL, true, true>(w0, w1, u, N, chunk, row, v, M, K, out,
                                                        s_tab);
  else if (mode & 1u)
    superpose_f64_body<SIGN, R, WF, 0, OBL, true, false>(w0, w1, u, N, chunk, 0u, v, M, K, out,
                                                         s_tab);
  else
    superpose_f64_body<SIGN, R, WF, 0, OBL, false, false>(w0, w1, u, N, chunk, 0u, v, M, K,
                                                          out, s_tab);
}

// Source layout on the device, for calls on device pointers (the host-pointer
// and plan paths scan the positions on the host with the same rules):
//   layout[0] bit 0: every z equals source 0's (== on values);
//   layout[0] bit 1: the sources form rows of L = layout[1] consecutive
//   sources that share the row's first x (L = the run of source 0's x, at
//   least 4; the last row may be shorter);
//   layout[1]: L.
__global__ __launch_bounds__(kBlock) void k_layout_init(uint32_t N, uint32_t allow,
                                                        uint32_t* __restrict__ layout) {
  if (threadIdx.x == 0) {
    layout[0] = allow;
    layout[1] = N;
  }
}

// L = the first n with x_n != x_0 (N if none): an atomic minimum, order-free.
__global__ __launch_bounds__(kBlock) void k_layout_run(const double* __restrict__ u, uint32_t N,
                                                       uint32_t* __restrict__ layout) {
  const double x0 = u[0];
  for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < N; n += gridDim.x * kBlock)
    if (u[3 * (size_t)n] != x0) {  // this thread's first hit is its smallest
      atomicMin(&layout[1], n);
      break;
    }
}

__global__ __launch_bounds__(kBlock) void k_layout_check(const double* __restrict__ u,
                                                         uint32_t N,
                                                         uint32_t* __restrict__ layout) {
  const uint32_t L = layout[1];
  const double z0 = u[2];
  bool planar = true, rows = L >= 4;
  for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < N; n += gridDim.x * kBlock) {
    planar = planar && u[3 * (size_t)n + 2] == z0;
    rows = rows && u[3 * (size_t)n] == u[3 * (size_t)(n - n % L)];
  }
  const uint32_t clear = (planar ? 0u : 3u) | (rows ? 0u : 2u);
  if (clear) atomicAnd(&layout[0], ~clear);
}

hipError_t launch_source_layout(const double* u, size_t N, uint32_t* layout,
                                hipStream_t stream, unsigned allow) {
  k_layout_init<<<1, kBlock, 0, stream>>>((uint32_t)N, allow, layout);
  if (!N) return hipGetLastError();
  const unsigned blocks = (unsigned)std::min<size_t>((N + kBlock - 1) / kBlock, 2048);
  k_layout_run<<<blocks, kBlock, 0, stream>>>(u, (uint32_t)N, layout);
  k_layout_check<<<blocks, kBlock, 0, stream>>>(u, (uint32_t)N, layout);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_reduce_partials(
    const double2* __restrict__ partial, uint32_t S, uint32_t M,
    double2* __restrict__ y) {
  const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= M) return;
  double re = 0.0, im = 0.0;
  for (uint32_t s = 0; s < S; ++s) {  // fixed order: deterministic
    const double2 p = partial[(size_t)s * M + m];
    re += p.x;
    im += p.y;
  }
  y[m] = make_double2(re, im);
}

// Few targets over many splits (the error report's 256 targets over up to
// 1024 splits): one workgroup per target, lane t summing the splits t, t + 256,
// ... in order, then a fixed tree over the lanes, so the order is set by S
// alone.
__global__ __launch_bounds__(kBlock) void k_reduce_partials_wide(
    const double2* __restrict__ partial, uint32_t S, uint32_t M,
    double2* __restrict__ y) {
  __shared__ double s_re[kBlock], s_im[kBlock];
  const uint32_t m = blockIdx.x, t = threadIdx.x;
  double re = 0.0, im = 0.0;
  for (uint32_t s = t; s < S; s += kBlock) {
    const double2 p = partial[(size_t)s * M + m];
    re += p.x;
    im += p.y;
  }
  s_re[t] = re;
  s_im[t] = im;
  __syncthreads();
  for (uint32_t h = kBlock / 2; h > 0; h >>= 1) {
    if (t < h) {
      s_re[t] += s_re[t + h];
      s_im[t] += s_im[t + h];
    }
    __syncthreads();
  }
  if (t == 0) y[m] = make_double2(s_re[0], s_im[0]);
}

hipError_t launch_reduce_partials(int S, size_t M, const double2* partial, double2* y,
                                  hipStream_t stream) {
  if (S <= 1 || M == 0) return hipSuccess;
  if (M <= 4096 && S >= 64) {
    k_reduce_partials_wide<<<(unsigned)M, kBlock, 0, stream>>>(partial, (uint32_t)S,
                                                               (uint32_t)M, y);
    return hipGetLastError();
  }
  const unsigned blocks = (unsigned)((M + kBlock - 1) / kBlock);
  k_reduce_partials<<<blocks, kBlock, 0, stream>>>(partial, (uint32_t)S,
                                                   (uint32_t)M, y);
  return hipGetLastError();
}

hipError_t launch_superpose_f64(int sign, WeightForm wf, bool obliquity, bool uniform_z,
                                size_t row, const double* w0, const double* w1, const double* u,
                                size_t N,
                                const double* v, size_t M, double K,
                                double2* y, double2* partial,
                                size_t partial_cap, hipStream_t stream, bool few_targets) {
  if (M == 0) return hipSuccess;
  int R = few_targets ? 1 : tuning().r_f64;
  if (R != 1 && R != 2 && R != 4) R = 2;
  int S = N ? choose_splits(M, R, N, 1024, few_targets ? kFewTargetSplits : 64) : 1;
  while (S > 1 && (size_t)S * M > partial_cap) --S;
  const uint32_t chunk = N ? (uint32_t)((N + S - 1) / S) : 0;
  const dim3 grid((unsigned)((M + (size_t)kBlock * R - 1) / ((size_t)kBlock * R)),
                  (unsigned)S);
  double2* out = S > 1 ? partial : y;
  with_sign(sign, [&](auto sg) {
    with_r(R, [&](auto rr) {
      with_bool(wf == kCartesian, [&](auto cart) {
        with_bool(obliquity, [&](auto obl) {
          constexpr int wfv = cart.value ? kCartesian : kPolar;
          if (tuning().f64_libm) {
            k_superpose_f64<sg.value, rr.value, wfv, 1, (bool)obl.value, false, false>
                <<<grid, kBlock, 0, stream>>>(w0, w1, u, (uint32_t)N, chunk, 0u, v,
                                              (uint32_t)M, K, out);
          } else if (uniform_z && tuning().f64_planar && row && tuning().f64_rows) {
            k_superpose_f64<sg.value, rr.value, wfv, 0, (bool)obl.value, true, true>
                <<<grid, kBlock, 0, stream>>>(w0, w1, u, (uint32_t)N, chunk, (uint32_t)row, v,
                                              (uint32_t)M, K, out);
          } else {
            with_bool(uniform_z && tuning().f64_planar, [&](auto pz) {
              k_superpose_f64<sg.value, rr.value, wfv, 0, (bool)obl.value, (bool)pz.value, false>
                  <<<grid, kBlock, 0, stream>>>(w0, w1, u, (uint32_t)N, chunk, 0u, v,
                                                (uint32_t)M, K, out);
            });
          }
        });
      });
    });
  });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_reduce_partials(S, M, partial, y, stream);
}

hipError_t launch_superpose_f64_auto(int sign, WeightForm wf, bool obliquity,
                                     const double* w0, const double* w1, const double* u,
                                     size_t N, const double* v, size_t M, double K, double2* y,
                                     double2* partial, size_t partial_cap, uint32_t* layout,
                                     hipStream_t stream) {
  if (M == 0) return hipSuccess;
  // the options switch the paths off as they do for host-scanned sources
  const unsigned allow = tuning().f64_planar ? (tuning().f64_rows ? 3u : 1u) : 0u;
  hipError_t e = launch_source_layout(u, N, layout, stream, allow);
  if (e != hipSuccess) return e;
  int R = tuning().r_f64;
  if (R != 1 && R != 2 && R != 4) R = 2;
  int S = N ? choose_splits(M, R, N, 1024) : 1;
  while (S > 1 && (size_t)S * M > partial_cap) --S;
  const uint32_t chunk = N ? (uint32_t)((N + S - 1) / S) : 0;
  const dim3 grid((unsigned)((M + (size_t)kBlock * R - 1) / ((size_t)kBlock * R)),
                  (unsigned)S);
  double2* out = S > 1 ? partial : y;
  with_sign(sign, [&](auto sg) {
    with_r(R, [&](auto rr) {
      with_bool(wf == kCartesian, [&](auto cart) {
        with_bool(obliquity, [&](auto obl) {
          constexpr int wfv = cart.value ? kCartesian : kPolar;
          k_superpose_f64_auto<sg.value, rr.value, wfv, (bool)obl.value>
              <<<grid, kBlock, 0, stream>>>(w0, w1, u, (uint32_t)N, chunk, layout, v,
                                            (uint32_t)M, K, out);
        });
      });
    });
  });
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_reduce_partials(S, M, partial, y, stream);
}

// ---------------------------------------------------------------------------
// FP32-split brute force over tiles of T sources.
//
// Per (target, tile), in FP64: r = v - c, d_ref = |r|, the fractional cycle
// f = frac(d_ref / lambda), and the tile rotation e^{i sign 2 pi f}.
// Per pair, in FP32 and without cancellation:
//   q = |delta|^2 - 2 r.delta          (= d^2 - d_ref^2, delta = u - c)
//   t = (d - d_ref) / lambda = q (alpha + c2 q)        to O(q^3), with
//       alpha = 1 / (2 d_ref lambda),  c2 = -1 / (8 d_ref^3 lambda)
//   1/d = 1/d_ref + q g,  g = -1 / (2 d_ref^3)          to O(q^2)
//   term = w * e^{i sign 2 pi t} / d
// (d = sqrt(d_ref^2 + q) = d_ref + q/(2 d_ref) - q^2/(8 d_ref^3) + ...).  The
// T terms of a tile are summed in FP32, rotated by the tile's FP64-derived
// phase and added into FP64 accumulators.  PLANAR: every slot of every tile
// has delta_z = 0 (sources on planes z = const), so the z term of q drops.
// Valid in the far field of each tile: |q| / d_ref^2 << 1 (tile radius much
// smaller than the distance) and |t| of order one cycle.

template <int SIGN, int R, int T, bool PLANAR, bool OBL>
__global__ __launch_bounds__(kBlock) void k_superpose_f32s(
    const double4* __restrict__ centre, const float4* __restrict__ geo,
    const float2* __restrict__ wc, const int32_t* __restrict__ perm,
    const double* __restrict__ u, uint32_t n_tiles, uint32_t chunk_tiles,
    const double* __restrict__ v, uint32_t M, double lambda,
    double inv_lambda, double2* __restrict__ out) {
  const uint32_t t0 = blockIdx.y * chunk_tiles;
  const uint32_t t1 = min(n_tiles, t0 + chunk_tiles);
  const uint32_t base = blockIdx.x * (kBlock * R) + threadIdx.x;
  double vx[R], vy[R], vz[R], ar[R], ai[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t m = base + r * kBlock;
    const size_t mm = m < M ? m : M - 1;
    vx[r] = v[3 * mm];
    vy[r] = v[3 * mm + 1];
    vz[r] = v[3 * mm + 2];
    ar[r] = 0.0;
    ai[r] = 0.0;
  }
  for (uint32_t t = t0; t < t1; ++t) {
    const double4 c = centre[t];
    const float2 rho = __builtin_bit_cast(float2, c.w);  // lateral, axial tile radius
    float m2rx[R], m2ry[R], m2rz[R], alpha[R], c2[R], gq[R], invd[R], rzf[R];
    float ct[R], st[R], sr[R], si[R];
    bool near[R];
    bool any_near = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double rx = vx[r] - c.x, ry = vy[r] - c.y, rz = vz[r] - c.z;
      double dref, id;
      hm_sqrt_inv(__builtin_fma(rz, rz, __builtin_fma(ry, ry, rx * rx)), dref, id);
      {
        const float rxy = __builtin_sqrtf((float)__builtin_fma(ry, ry, rx * rx));
        const float qmax = 2.f * (rxy * rho.x + __builtin_fabsf((float)rz) * rho.y) +
                           rho.x * rho.x + rho.y * rho.y;
        const float ia = (float)(id * inv_lambda);  // 1 / (d lambda)
        const float idf = (float)id;
        const float round_err = 3e-8f * qmax * ia;
        const float trunc_err = 0.0625f * qmax * qmax * qmax * ia * idf * idf * idf * idf;
        // NaN (a target at the tile centre) also counts as near
        near[r] = !(round_err <= kGuardCycles && trunc_err <= 0.05f * kGuardCycles);
        any_near |= near[r];
      }
      const float f = (float)frac_cycle(dref * inv_lambda);
      sincos_rev_f32(f, st[r], ct[r]);
      m2rx[r] = (float)(-2.0 * rx);
      m2ry[r] = (float)(-2.0 * ry);
      m2rz[r] = (float)(-2.0 * rz);
      const double id2 = id * id;
      const double a = 0.5 * id * inv_lambda;
      alpha[r] = (float)a;
      c2[r] = (float)(-0.25 * a * id2);
      if constexpr (OBL) {
        // |dz| / d^2 = |dz| (1/d_ref^2 - q/d_ref^4 + O(q^2)); planar tiles
        // have dz = rz for every slot, so |rz| folds into the coefficients.
        const double zs = PLANAR ? __builtin_fabs(rz) : 1.0;
        invd[r] = (float)(zs * id2);
        gq[r] = (float)(-zs * id2 * id2);
        rzf[r] = (float)rz;
      } else {
        gq[r] = (float)(-0.5 * id2 * id);
        invd[r] = (float)id;
        rzf[r] = 0.f;
      }
      sr[r] = 0.f;
      si[r] = 0.f;
    }
    const float4* g = geo + (size_t)t * T;
    const float2* w = wc + (size_t)t * T;
    // a wave whose every target is near this tile skips the expansion
    bool all_near = true;
#pragma unroll
    for (int r = 0; r < R; ++r) all_near &= near[r];
    if (__builtin_amdgcn_ballot_w64(!all_near) != 0) {
#pragma unroll 4
      for (int j = 0; j < T; ++j) {
        const float4 gj = g[j];
        const float2 wj = w[j];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          float q = __builtin_fmaf(m2ry[r], gj.y, gj.w);
          if constexpr (!PLANAR) q = __builtin_fmaf(m2rz[r], gj.z, q);
          q = __builtin_fmaf(m2rx[r], gj.x, q);
          // w e^{i s 2 pi t} = |w| e^{2 pi i (arg w + s t)}: the weight's phase
          // joins the pair's, its modulus the amplitude
          const float pc = __builtin_fmaf(c2[r], q, alpha[r]);
          const float tt = SIGN > 0 ? __builtin_fmaf(q, pc, wj.y) : __builtin_fmaf(-q, pc, wj.y);
          float sn, cs;
          sincos_rev_f32(tt, sn, cs);
          float amp = wj.x * __builtin_fmaf(q, gq[r], invd[r]);
          if constexpr (OBL && !PLANAR) amp *= __builtin_fabsf(rzf[r] - gj.z);
          sr[r] = __builtin_fmaf(amp, cs, sr[r]);
          si[r] = __builtin_fmaf(amp, sn, si[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double xr = near[r] ? 0.0 : sr[r], xi = near[r] ? 0.0 : si[r];
      const double cr = ct[r], cs = SIGN > 0 ? (double)st[r] : -(double)st[r];
      ar[r] += xr * cr - xi * cs;
      ai[r] += xr * cs + xi * cr;
    }
    if (any_near)  // thick tiles or near field
      exact_tile_pairs_polar<SIGN, OBL, R, PLANAR>(u, perm + (size_t)t * T, w, T, c.z, vx, vy,
                                                   vz, near, inv_lambda, ar, ai);
  }
  double2* dst = out + (size_t)blockIdx.y * M;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t m = base + r * kBlock;
    if (m < M) dst[m] = make_double2(ar[r], ai[r]);
  }
}

template <typename F>
static bool with_tile(unsigned T, F&& f) {
  switch (T) {
    case 8: f(IC<8>{}); return true;
    case 16: f(IC<16>{}); return true;
    case 32: f(IC<32>{}); return true;
    case 64: f(IC<64>{}); return true;
    default: return false;
  }
}

hipError_t launch_superpose_f32s(int sign, bool obliquity, const TileSet& ts,
                                 const float2* wc, const double* v, size_t M,
                                 double lambda, double2* y, double2* partial,
                                 size_t partial_cap, hipStream_t stream) {
  if (M == 0) return hipSuccess;
  int R = tuning().r_f32;
  if (R != 1 && R != 2 && R != 4) R = 2;
  int S = ts.n_tiles ? choose_splits(M, R, ts.n_tiles, 64) : 1;
  while (S > 1 && (size_t)S * M > partial_cap) --S;
  const uint32_t chunk = ts.n_tiles ? (ts.n_tiles + S - 1) / S : 0;
  const dim3 grid((unsigned)((M + (size_t)kBlock * R - 1) / ((size_t)kBlock * R)),
                  (unsigned)S);
  double2* out = S > 1 ? partial : y;
  const double il = 1.0 / lambda;
  const bool ok = with_tile(ts.tile, [&](auto tt) {
    with_sign(sign, [&](auto sg) {
      with_r(R, [&](auto rr) {
        with_bool(ts.planar, [&](auto pl) {
          with_bool(obliquity, [&](auto obl) {
            k_superpose_f32s<sg.value, rr.value, tt.value, (bool)pl.value, (bool)obl.value>
                <<<grid, kBlock, 0, stream>>>(ts.centre, ts.geo, wc, ts.perm, ts.u, ts.n_tiles,
                                              chunk, v, (uint32_t)M, lambda, il, out);
          });
        });
      });
    });
  });
  if (!ok) return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_reduce_partials(S, M, partial, y, stream);
}

__global__ __launch_bounds__(kBlock) void k_tile_weights(
    int wf, const double* __restrict__ w0, const double* __restrict__ w1,
    const int32_t* __restrict__ perm, uint32_t n_slots,
    float2* __restrict__ wc, int polar_out) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n_slots) return;
  const int32_t n = perm[s];
  float2 r = make_float2(0.f, 0.f);
  if (n >= 0) {
    constexpr double kInv2Pi = 0.15915494309189535;
    if (polar_out) {  // (|w|, arg w in cycles, reduced to [-1/2, 1/2))
      double a, t;
      if (wf == kPolar) {
        a = w0[n];
        t = w1[n] * kInv2Pi + (a < 0.0 ? 0.5 : 0.0);
        a = __builtin_fabs(a);
      } else {
        const double re = w0[2 * (size_t)n], im = w0[2 * (size_t)n + 1];
        a = hypot(re, im);
        t = atan2(im, re) * kInv2Pi;
      }
      r = make_float2((float)a, (float)(t - __builtin_rint(t)));
    } else if (wf == kPolar) {
      double sn, cs;
      sincos(w1[n], &sn, &cs);
      r = make_float2((float)(w0[n] * cs), (float)(w0[n] * sn));
    } else {
      r = make_float2((float)w0[2 * (size_t)n], (float)w0[2 * (size_t)n + 1]);
    }
  }
  wc[s] = r;
}

hipError_t launch_tile_weights(WeightForm wf, const double* w0,
                               const double* w1, const int32_t* perm,
                               size_t n_slots, float2* wc,
                               hipStream_t stream, bool polar_out) {
  if (!n_slots) return hipSuccess;
  const unsigned blocks = (unsigned)((n_slots + kBlock - 1) / kBlock);
  k_tile_weights<<<blocks, kBlock, 0, stream>>>((int)wf, w0, w1, perm,
                                               (uint32_t)n_slots, wc, polar_out ? 1 : 0);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Monte Carlo: target m draws K indices n_{m,k} from Philox4x32-10 keyed by
// the seed with counter (k/2, m, 0, kMcStream); words (x,y) serve even k,
// (z,w) odd k; Lemire multiply-shift maps 64 bits into [0, N).
// Batched ("shared") Monte Carlo: the targets of group g = m / kMcGroup share
// one subset, drawn with counter (k/2, g, kMcSharedTag, kMcStream); the
// index is then wave-uniform, Philox runs on the scalar unit and sources
// arrive through scalar loads.

constexpr uint32_t kMcSharedTag = 0x80000000u;

// The Philox block of draws k, k + 1 (k even) of target m: per target, or
// (SHARED_IDX) of m's group's shared subset.
template <bool SHARED_IDX>
__device__ __forceinline__ U4 mc_draws(uint32_t k, uint32_t m, uint32_t slo, uint32_t shi) {
  if constexpr (SHARED_IDX)
    return philox4x32_10(U4{k >> 1, m / kMcGroup, kMcSharedTag, kMcStream}, slo, shi);
  else
    return philox4x32_10(U4{k >> 1, m, 0u, kMcStream}, slo, shi);
}

// A drawn source n as (x, y, z, w): w = amp (FP64, the phase in ph[n]) or the
// bit-cast FP32 weight (FP32-split).  On a product grid (FP32-split only) the
// sources of the leading complete rows come from the row and column tables
// and the weight array, the same values as the packed records.
struct McSourceView {
  const double4* pos;
  McLattice lat;
  uint32_t ny = 0, body = 0;
  double inv_ny = 0.0;
  __device__ McSourceView(const double4* p, const McLattice& l, bool fp64, uint32_t N)
      : pos(p), lat(l) {
    uint32_t rows = 0;
    if (!fp64 && lat.info) {
      ny = lat.info[0];
      rows = ny ? lat.info[1] : 0u;
      if (!rows) ny = 0;
    }
    body = min(rows * ny, N);
    inv_ny = ny ? 1.0 / (double)ny : 0.0;
  }
  // branch-free: each component is one load from an address chosen by select
  __device__ __forceinline__ double4 operator()(uint32_t n) const {
    const bool tab = n < body;
    uint32_t i = (uint32_t)((double)n * inv_ny);  // within one of n / ny
    i -= (uint32_t)(i * ny > n);
    i += (uint32_t)((i + 1) * ny <= n);
    const uint32_t j = n - i * ny;
    const double* rec = &pos[n].x;
    const double* px = tab ? lat.xy + kLatticeMaxRow + 1 + i : rec;
    const double* py = tab ? lat.xy + j : rec + 1;
    const double* pz = tab ? lat.xy + kLatticeMaxRow : rec + 2;
    const double* pw = tab ? reinterpret_cast<const double*>(lat.wc + n) : rec + 3;
    return make_double4(*px, *py, *pz, *pw);
  }
};

// Per-target estimator; thread i evaluates the target at v[i], whose index
// (the Philox counter) is tmap[i], or i without a map.  Block row blockIdx.y
// takes the draws k in [blockIdx.y kchunk, +kchunk) (kchunk even) of the K and
// writes its share to y[blockIdx.y M + i].
template <int SIGN, bool FP64, bool OBL, bool SHARED_IDX>
__global__ __launch_bounds__(kBlock) void k_mc(
    const double4* __restrict__ pos, const double* __restrict__ ph,
    uint32_t N, const double* __restrict__ v, uint32_t M, const uint32_t* __restrict__ tmap,
    uint32_t K, uint32_t kchunk, uint32_t seed_lo, uint32_t seed_hi, double Kw,
    double inv_lambda, double scale, McLattice lat, double2* __restrict__ y) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= M) return;
  const uint32_t kb = blockIdx.y * kchunk, ke = min(K, kb + kchunk);
  const uint32_t m = tmap ? tmap[i] : i;
  const double vx = v[3 * (size_t)i], vy = v[3 * (size_t)i + 1],
               vz = v[3 * (size_t)i + 2];
  const McSourceView source(pos, lat, FP64, N);
  double ar = 0.0, ai = 0.0;
  float fr = 0.f, fi = 0.f;
  // 8 samples per step: 4 Philox blocks, then 8 independent gathers in flight
  const uint32_t K8 = kb + ((ke - kb) & ~7u);
  for (uint32_t k = kb; k < K8; k += 8) {
    uint32_t idx[8];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const U4 o = mc_draws<SHARED_IDX>(k + 2 * h, m, seed_lo, seed_hi);
      idx[2 * h] = lemire_index(o.x, o.y, N);
      idx[2 * h + 1] = lemire_index(o.z, o.w, N);
    }
    double4 p[8];
    double phv[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      p[h] = source(idx[h]);
      if constexpr (FP64) phv[h] = ph[idx[h]];
    }
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      if constexpr (FP64) {
        const PairF64 q =
            pair_f64<SIGN, 0, OBL>(p[h].w, phv[h], p[h].x, p[h].y, p[h].z, vx, vy, vz, Kw);
        accumulate_phasor<0>(q.r, q.arg, ar, ai);
      } else {
        mc_pair_f32s<SIGN, OBL>(p[h], vx, vy, vz, inv_lambda, fr, fi);
      }
    }
    if constexpr (!FP64) {
      ar += fr;
      ai += fi;
      fr = 0.f;
      fi = 0.f;
    }
  }
  for (uint32_t k = K8; k < ke; k += 2) {
    const U4 o = mc_draws<SHARED_IDX>(k, m, seed_lo, seed_hi);
    const uint32_t idx[2] = {lemire_index(o.x, o.y, N), lemire_index(o.z, o.w, N)};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && k + 1 >= ke) break;
      const double4 p = source(idx[h]);
      if constexpr (FP64) {
        const PairF64 q = pair_f64<SIGN, 0, OBL>(p.w, ph[idx[h]], p.x, p.y, p.z, vx, vy, vz, Kw);
        accumulate_phasor<0>(q.r, q.arg, ar, ai);
      } else {
        mc_pair_f32s<SIGN, OBL>(p, vx, vy, vz, inv_lambda, fr, fi);
      }
    }
  }
  ar += fr;
  ai += fi;
  y[(size_t)blockIdx.y * M + i] = make_double2(ar * scale, ai * scale);
}

// Per-target Monte Carlo over source ranges.  A per-target draw gathers one
// random source; from a 1024^2 plane the gathered data (8 MB of weights on a
// lattice, 32 MB of records otherwise) does not fit the 4 MB L2 of an XCD, and
// the misses bound k_mc.  k_mc_pass takes only the draws whose source index
// lies in [lo, hi), a range whose data fits one L2, and the host launches it
// once per range: every launch regenerates its targets' Philox streams (the
// indices are k_mc's), queues each lane's in-range draws in that lane's LDS
// ring and evaluates them 8 at a time, every lane of the wave taking 8 once
// each active lane holds 8; a ring that could overflow in the next round
// forces a partial batch (lanes with fewer entries idle in it).  Each launch
// adds its sums into y[m] (FP64): the first writes, the last scales by N / K.
// The summation order of a target is fixed by its indices and the ranges.
constexpr uint32_t kMcQueue = 32;  // ring slots per lane, a power of two

template <int SIGN, bool FP64, bool OBL>
__global__ __launch_bounds__(kBlock) void k_mc_pass(
    const double4* __restrict__ pos, const double* __restrict__ ph, uint32_t N,
    const double* __restrict__ v, uint32_t M, uint32_t K, uint32_t seed_lo, uint32_t seed_hi,
    double Kw, double inv_lambda, McLattice lat, uint32_t lo, uint32_t hi, int first,
    double scale, double2* __restrict__ y) {
  __shared__ uint32_t s_q[kMcQueue][kBlock];  // [slot][lane]: one bank per lane
  const uint32_t t = threadIdx.x, m = blockIdx.x * kBlock + t;
  const bool active = m < M;
  if (__ballot(active) == 0) return;  // the whole wave is past M (no barriers below)
  const uint32_t mm = active ? m : M - 1;
  const double vx = v[3 * (size_t)mm], vy = v[3 * (size_t)mm + 1], vz = v[3 * (size_t)mm + 2];
  const McSourceView source(pos, lat, FP64, N);
  double ar = 0.0, ai = 0.0;
  uint32_t head = 0, tail = 0;
  // the next `take` (<= 8) queued draws of this lane; idle slots read source
  // lo and their terms are dropped by a select (never multiplied by zero: a
  // target at that source would turn 0 * inf into NaN)
  const auto batch = [&](uint32_t take) __attribute__((always_inline)) {
    uint32_t idx[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      const uint32_t q = s_q[(head + h) & (kMcQueue - 1)][t];
      idx[h] = (active && (uint32_t)h < take) ? q : lo;
    }
    double4 p[8];
    double phv[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      p[h] = source(idx[h]);
      if constexpr (FP64) phv[h] = ph[idx[h]];
    }
    float fr = 0.f, fi = 0.f;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      const bool use = (uint32_t)h < take;
      if constexpr (FP64) {
        const PairF64 q =
            pair_f64<SIGN, 0, OBL>(p[h].w, phv[h], p[h].x, p[h].y, p[h].z, vx, vy, vz, Kw);
        double tr = 0.0, ti = 0.0;
        accumulate_phasor<0>(q.r, q.arg, tr, ti);
        ar += use ? tr : 0.0;
        ai += use ? ti : 0.0;
      } else {
        float tr = 0.f, ti = 0.f;
        mc_pair_f32s<SIGN, OBL>(p[h], vx, vy, vz, inv_lambda, tr, ti);
        fr += use ? tr : 0.f;
        fi += use ? ti : 0.f;
      }
    }
    if constexpr (!FP64) {
      ar += fr;
      ai += fi;
    }
    head += take;
  };
  for (uint32_t k = 0; k < K; k += 8) {
    uint32_t idx[8];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const U4 o = mc_draws<false>(k + 2 * h, mm, seed_lo, seed_hi);
      idx[2 * h] = lemire_index(o.x, o.y, N);
      idx[2 * h + 1] = lemire_index(o.z, o.w, N);
    }
#pragma unroll
    for (int h = 0; h < 8; ++h)
      if (active && k + h < K && idx[h] - lo < hi - lo) {
        s_q[tail & (kMcQueue - 1)][t] = idx[h];
        ++tail;
      }
    // full batches while every active lane holds 8 (idle lanes count as full)
    while (__builtin_amdgcn_ballot_w64(active && tail - head < 8) == 0) batch(8);
    // every ring holds at most kMcQueue - 8 entries before the next round
    if (__builtin_amdgcn_ballot_w64(active && tail - head > kMcQueue - 8))
      batch(min(8u, tail - head));
  }
  while (__builtin_amdgcn_ballot_w64(active && tail != head)) batch(min(8u, tail - head));
  if (active) {
    if (!first) {
      const double2 old = y[m];
      ar += old.x;
      ai += old.y;
    }
    y[m] = make_double2(ar * scale, ai * scale);
  }
}

// Batched Monte Carlo: one workgroup = one group of kMcGroup targets
// (kBlock lanes x 4 targets each) sharing the subset n_{g,k}.
template <int SIGN, bool FP64, bool OBL>
__global__ __launch_bounds__(kBlock) void k_mc_shared(
    const double4* __restrict__ pos, const double* __restrict__ ph,
    uint32_t N, const double* __restrict__ v, uint32_t M, uint32_t K,
    uint32_t seed_lo, uint32_t seed_hi, double Kw, double inv_lambda,
    double scale, double2* __restrict__ y) {
  constexpr int R = kMcGroup / kBlock;
  const uint32_t g = blockIdx.x;
  double vx[R], vy[R], vz[R], ar[R], ai[R];
  float fr[R], fi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t m = g * kMcGroup + r * kBlock + threadIdx.x;
    const size_t mm = m < M ? m : M - 1;
    vx[r] = v[3 * mm];
    vy[r] = v[3 * mm + 1];
    vz[r] = v[3 * mm + 2];
    ar[r] = ai[r] = 0.0;
    fr[r] = fi[r] = 0.f;
  }
  for (uint32_t k = 0; k < K; k += 2) {
    const U4 o = philox4x32_10(U4{k >> 1, g, kMcSharedTag, kMcStream}, seed_lo, seed_hi);
    const uint32_t idx[2] = {lemire_index(o.x, o.y, N), lemire_index(o.z, o.w, N)};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && k + 1 >= K) break;
      const double4 p = pos[idx[h]];
      if constexpr (FP64) {
        const double phv = ph[idx[h]];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const PairF64 q =
              pair_f64<SIGN, 0, OBL>(p.w, phv, p.x, p.y, p.z, vx[r], vy[r], vz[r], Kw);
          accumulate_phasor<0>(q.r, q.arg, ar[r], ai[r]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
          mc_pair_f32s<SIGN, OBL>(p, vx[r], vy[r], vz[r], inv_lambda, fr[r], fi[r]);
      }
    }
    if constexpr (!FP64) {
      if ((k & 62) == 62 || k + 2 >= K) {  // flush FP32 partials every 64 draws
#pragma unroll
        for (int r = 0; r < R; ++r) {
          ar[r] += fr[r];
          ai[r] += fi[r];
          fr[r] = fi[r] = 0.f;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t m = g * kMcGroup + r * kBlock + threadIdx.x;
    if (m < M) y[m] = make_double2(ar[r] * scale, ai[r] * scale);
  }
}

uint32_t mc_passes(size_t N, bool fp64) {
  const int opt = tuning().mc_passes;
  if (opt > 0) return (uint32_t)std::min<size_t>((size_t)opt, std::max<size_t>(N, 1));
  // the gathered bytes per source: the FP32 weight on a lattice (the records
  // of other point sets are 4x larger and get 4x fuller ranges), or the FP64
  // record and phase
  const size_t bytes = N * (fp64 ? 40 : 8);
  const size_t p = (bytes + kMcPassBytes - 1) / kMcPassBytes;
  return (uint32_t)std::min<size_t>(std::max<size_t>(p, 1), 8);
}

hipError_t launch_mc(int sign, bool fp64, bool shared, bool obliquity, const double4* pos,
                     const double* ph, size_t N, const double* v, size_t M,
                     uint32_t K, uint64_t seed, double K_wave, double lambda,
                     const McLattice& lat, double2* y, hipStream_t stream) {
  if (!M) return hipSuccess;
  const double scale = (double)N / (double)K;
  const uint32_t slo = (uint32_t)seed, shi = (uint32_t)(seed >> 32);
  const double il = 1.0 / lambda;
  const unsigned blocks = shared ? (unsigned)((M + kMcGroup - 1) / kMcGroup)
                                 : (unsigned)((M + kBlock - 1) / kBlock);
  const uint32_t passes = shared ? 1u : mc_passes(N, fp64);
  with_sign(sign, [&](auto sg) {
    with_bool(fp64, [&](auto f64) {
      with_bool(obliquity, [&](auto obl) {
        if (shared)
          k_mc_shared<sg.value, (bool)f64.value, (bool)obl.value><<<blocks, kBlock, 0, stream>>>(
              pos, ph, (uint32_t)N, v, (uint32_t)M, K, slo, shi, K_wave, il, scale, y);
        else if (passes <= 1)
          k_mc<sg.value, (bool)f64.value, (bool)obl.value, false>
              <<<blocks, kBlock, 0, stream>>>(pos, ph, (uint32_t)N, v, (uint32_t)M, nullptr, K,
                                              K + (K & 1u), slo, shi, K_wave, il, scale, lat, y);
        else
          for (uint32_t b = 0; b < passes; ++b)
            k_mc_pass<sg.value, (bool)f64.value, (bool)obl.value>
                <<<blocks, kBlock, 0, stream>>>(
                    pos, ph, (uint32_t)N, v, (uint32_t)M, K, slo, shi, K_wave, il, lat,
                    (uint32_t)((uint64_t)b * N / passes),
                    (uint32_t)((uint64_t)(b + 1) * N / passes), b == 0 ? 1 : 0,
                    b + 1 == passes ? scale : 1.0, y);
      });
    });
  });
  return hipGetLastError();
}

hipError_t launch_mc_subset(int sign, bool shared, bool obliquity, const double4* pos,
                            const double* ph, size_t N, const double* vsub, size_t count,
                            const uint32_t* tmap, uint32_t K, uint64_t seed, double K_wave,
                            double lambda, double2* y, double2* partial, size_t partial_cap,
                            hipStream_t stream) {
  if (!count) return hipSuccess;
  const double scale = (double)N / (double)K;
  const uint32_t slo = (uint32_t)seed, shi = (uint32_t)(seed >> 32);
  const unsigned blocks = (unsigned)((count + kBlock - 1) / kBlock);
  // a few hundred targets: the draws are split into chunks of at least 64
  // (a multiple of 8) over enough blocks to fill the GPU, summed in order
  uint32_t S = (uint32_t)std::min<size_t>({(K + 63) / 64, kFewTargetSplits,
                                          partial_cap / count});
  S = std::max(S, 1u);
  uint32_t chunk = (K + S - 1) / S;
  chunk = (chunk + 7u) & ~7u;
  S = (K + chunk - 1) / chunk;
  double2* out = S > 1 ? partial : y;
  with_sign(sign, [&](auto sg) {
    with_bool(shared, [&](auto sh) {
      with_bool(obliquity, [&](auto obl) {
        k_mc<sg.value, true, (bool)obl.value, (bool)sh.value>
            <<<dim3(blocks, S), kBlock, 0, stream>>>(pos, ph, (uint32_t)N, vsub,
                                                     (uint32_t)count, tmap, K, chunk, slo, shi,
                                                     K_wave, 1.0 / lambda, scale, McLattice{},
                                                     out);
      });
    });
  });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_reduce_partials((int)S, count, partial, y, stream);
}

__global__ __launch_bounds__(kBlock) void k_mc_pack(
    int fp64, const double* __restrict__ amp, const double* __restrict__ phase,
    const double* __restrict__ u, uint32_t N, double4* __restrict__ pos,
    double* __restrict__ ph, float2* __restrict__ wc) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double4 p;
  p.x = u[3 * (size_t)n];
  p.y = u[3 * (size_t)n + 1];
  p.z = u[3 * (size_t)n + 2];
  if (fp64) {
    p.w = amp[n];
    ph[n] = phase[n];
  } else {
    double sn, cs;
    sincos(phase[n], &sn, &cs);
    const float2 w = make_float2((float)(amp[n] * cs), (float)(amp[n] * sn));
    p.w = __builtin_bit_cast(double, w);
    if (wc) wc[n] = w;
  }
  pos[n] = p;
}

hipError_t launch_mc_pack(bool fp64, const double* amp, const double* phase,
                          const double* u, size_t N, double4* pos, double* ph, float2* wc,
                          hipStream_t stream) {
  if (!N) return hipSuccess;
  const unsigned blocks = (unsigned)((N + kBlock - 1) / kBlock);
  k_mc_pack<<<blocks, kBlock, 0, stream>>>(fp64 ? 1 : 0, amp, phase, u,
                                          (uint32_t)N, pos, ph, wc);
  return hipGetLastError();
}

// Row length of an x-slow product grid: the first n with x_n != x_0, among
// n < min(N, kLatticeMaxRow); N when all x are equal and N <= kLatticeMaxRow;
// 0 (not a grid) otherwise.  info[0] = that length ny, info[1] = N / ny, the
// number of whole rows, which k_lattice_check lowers to the number of leading
// complete rows.
__global__ __launch_bounds__(kBlock) void k_lattice_row(const double* __restrict__ u,
                                                        uint32_t N, uint32_t* __restrict__ info) {
  __shared__ uint32_t s_min[kBlock / 64];
  const double x0 = u[0];
  const uint32_t lim = min(N, kLatticeMaxRow);
  uint32_t first = lim;
  for (uint32_t n = 1 + threadIdx.x; n < lim; n += kBlock)
    if (u[3 * (size_t)n] != x0) {  // this thread's first hit is its smallest
      first = n;
      break;
    }
  for (int off = 32; off; off >>= 1) first = min(first, (uint32_t)__shfl_xor((int)first, off, 64));
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = first;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t f = lim;
    for (int w = 0; w < kBlock / 64; ++w) f = min(f, s_min[w]);
    const uint32_t ny = f < lim ? f : (N <= kLatticeMaxRow ? N : 0u);
    info[0] = ny;
    info[1] = ny ? N / ny : 0u;
  }
}

// Checks u[n] == (x of row n / ny, y of column n % ny, z_0) exactly for the
// sources of the whole rows and lowers info[1] to the first row where it fails
// (an atomic minimum: order-free), so that info[1] rows lead complete (a
// padded plane's copies of source 0 repeat row 0's y and never complete a
// row); fills the tables described with McLattice.
__global__ __launch_bounds__(kBlock) void k_lattice_check(const double* __restrict__ u,
                                                          uint32_t N, uint32_t* __restrict__ info,
                                                          double* __restrict__ xy) {
  const uint32_t ny = info[0];
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (!ny || n >= (N / ny) * ny) return;
  const uint32_t i = n / ny, j = n - i * ny;
  const double x = u[3 * (size_t)n], y = u[3 * (size_t)n + 1], z = u[3 * (size_t)n + 2];
  if (!(x == u[3 * (size_t)i * ny] && y == u[3 * (size_t)j + 1] && z == u[2]))
    atomicMin(&info[1], i);
  if (j == 0) xy[kLatticeMaxRow + 1 + i] = x;
  if (i == 0) xy[j] = y;
  if (n == 0) xy[kLatticeMaxRow] = z;
}

hipError_t launch_mc_lattice_probe(const double* u, size_t N, uint32_t* info, double* xy,
                                   hipStream_t stream) {
  if (!N) return hipSuccess;
  k_lattice_row<<<1, kBlock, 0, stream>>>(u, (uint32_t)N, info);
  const unsigned blocks = (unsigned)((N + kBlock - 1) / kBlock);
  k_lattice_check<<<blocks, kBlock, 0, stream>>>(u, (uint32_t)N, info, xy);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_mc_indices(
    uint32_t seed_lo, uint32_t seed_hi, uint64_t N, uint32_t m0,
    uint32_t count, uint32_t K, uint32_t tag, int64_t* __restrict__ idx) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= count) return;
  const uint32_t m = m0 + j;
  for (uint32_t k = 0; k < K; k += 2) {
    const U4 o = philox4x32_10(U4{k >> 1, m, tag, kMcStream}, seed_lo, seed_hi);
    idx[(size_t)j * K + k] = lemire_index(o.x, o.y, (uint32_t)N);
    if (k + 1 < K) idx[(size_t)j * K + k + 1] = lemire_index(o.z, o.w, (uint32_t)N);
  }
}

hipError_t launch_mc_indices(uint64_t seed, uint64_t N, size_t m0,
                             size_t count, uint32_t K, bool shared, int64_t* idx,
                             hipStream_t stream) {
  if (!count || !K) return hipSuccess;
  const unsigned blocks = (unsigned)((count + kBlock - 1) / kBlock);
  k_mc_indices<<<blocks, kBlock, 0, stream>>>((uint32_t)seed,
                                             (uint32_t)(seed >> 32), N,
                                             (uint32_t)m0, (uint32_t)count, K,
                                             shared ? kMcSharedTag : 0u, idx);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// time_transform post-processing (cuda/transform.cu:106-139, 400-405):
// max |y| (order-free, so deterministic), then polar(y / max).

constexpr unsigned kMaxBlocks = 1024;

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off));
  return x;
}

__global__ __launch_bounds__(kBlock) void k_max_abs(
    const double2* __restrict__ y, uint32_t M, double* __restrict__ block_max) {
  __shared__ double red[kBlock / 64];
  double mx = 0.0;
  for (uint32_t m = blockIdx.x * kBlock + threadIdx.x; m < M;
       m += gridDim.x * kBlock) {
    const double2 z = y[m];
    mx = fmax(mx, hypot(z.x, z.y));
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int i = 1; i < kBlock / 64; ++i) b = fmax(b, red[i]);
    block_max[blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(kBlock) void k_final_max(
    const double* __restrict__ block_max, uint32_t nb,
    double* __restrict__ max_out) {
  __shared__ double red[kBlock / 64];
  double mx = 0.0;
  for (uint32_t i = threadIdx.x; i < nb; i += kBlock) mx = fmax(mx, block_max[i]);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int i = 1; i < kBlock / 64; ++i) b = fmax(b, red[i]);
    max_out[0] = b;
  }
}

__global__ __launch_bounds__(kBlock) void k_to_polar(
    const double2* __restrict__ y, uint32_t M,
    const double* __restrict__ max_in, double* __restrict__ amp,
    double* __restrict__ phase) {
  const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= M) return;
  const double mx = max_in[0];
  const double div = mx < 1e-6 ? 1.0 : mx;
  const double2 z = y[m];
  const double re = z.x / div, im = z.y / div;
  amp[m] = hypot(re, im);
  phase[m] = atan2(im, re);
}

hipError_t launch_normalize_polar(const double2* y, size_t M, double* amp,
                                  double* phase, double* max_out,
                                  double* scratch, hipStream_t stream) {
  if (!M) return hipSuccess;
  unsigned nb = (unsigned)std::min<size_t>((M + kBlock - 1) / kBlock, kMaxBlocks);
  k_max_abs<<<nb, kBlock, 0, stream>>>(y, (uint32_t)M, scratch);
  k_final_max<<<1, kBlock, 0, stream>>>(scratch, nb, max_out);
  const unsigned blocks = (unsigned)((M + kBlock - 1) / kBlock);
  k_to_polar<<<blocks, kBlock, 0, stream>>>(y, (uint32_t)M, max_out, amp, phase);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Error report of an FP32-split result y (M values), from count FP64 values
// at sampled targets: n_e of them drawn by the result's energy and n_u
// spread evenly (multiple importance sampling with the balance heuristic).
// Energy draw i < n_e is the target at which the running sum of |y|^2 in
// target order first exceeds (i + 1/2) / n_e of the total T; even draw j is
// target floor(j M / n_u).  Then
//   err^2 = sum_i |y[m_i] - ref_i|^2 / (n_e |y[m_i]|^2 + n_u T / M),
// an unbiased estimate (under random offsets of the strata) of
// sum_m |y_m - y64_m|^2 / T, the whole result's squared relative L2 error:
// the energy draws find error that grows with the field (a focus), the even
// ones error spread over dark targets.  M <= count takes every target once
// (n_e = 0: the exact value).  A result without a finite positive T takes
// the even targets and sqrt(sum |y - ref|^2 / sum |ref|^2) there.  Every sum
// has a fixed order.

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

// inclusive prefix sum over the 64 lanes (a fixed order)
__device__ __forceinline__ double wave_scan(double x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double t = __shfl_up(x, off);
    if (lane >= off) x += t;
  }
  return x;
}

__device__ __forceinline__ double energy(const double2 a) { return a.x * a.x + a.y * a.y; }

// sum of E[0 .. n) by one wave, the same bits wherever it is called
__device__ __forceinline__ double energy_total(const double* __restrict__ E, uint32_t n) {
  double s = 0.0;
  for (uint32_t b = threadIdx.x & 63; b < n; b += 64) s += E[b];
  return wave_sum(s);
}

__device__ __forceinline__ bool energy_mode(double total) {
  return total > 0.0 && __builtin_isfinite(total);  // NaN fails
}

// Block b: E[b] = sum of |y|^2 over its chunk [b L, min(M, (b + 1) L)).
__global__ __launch_bounds__(kBlock) void k_energy_chunks(const double2* __restrict__ y,
                                                          uint32_t M, uint32_t L,
                                                          double* __restrict__ E) {
  __shared__ double red[kBlock / 64];
  const uint64_t b0 = (uint64_t)blockIdx.x * L, b1 = min<uint64_t>(M, b0 + L);
  double s = 0.0;
  for (uint64_t m = b0 + threadIdx.x; m < b1; m += kBlock) s += energy(y[m]);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += red[w];
    E[blockIdx.x] = t;
  }
}

// One wave per sample i: m_i as above, into tmap[i] and its position into vsub.
__global__ __launch_bounds__(kBlock) void k_error_targets(
    const double2* __restrict__ y, uint32_t M, uint32_t L, const double* __restrict__ E,
    uint32_t nchunks, const double* __restrict__ v, uint32_t count, uint32_t ne,
    uint32_t* __restrict__ tmap, double* __restrict__ vsub) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (i >= count) return;  // the whole wave
  const double total = energy_total(E, nchunks);
  const bool draw = energy_mode(total) && i < ne;
  const uint32_t j = energy_mode(total) ? i - ne : i, nu = energy_mode(total) ? count - ne : count;
  uint32_t m = draw ? 0u : (uint32_t)(((uint64_t)j * M) / nu);
  if (draw) {
    const double c = ((double)i + 0.5) / (double)ne * total;
    // the chunk where the running sum passes c (else the last one with energy)
    uint32_t cb = nchunks, lastb = nchunks;
    double before = 0.0, lastbefore = 0.0, run = 0.0;
    for (uint32_t base = 0; base < nchunks && cb == nchunks; base += 64) {
      const double e = base + lane < nchunks ? E[base + lane] : 0.0;
      const double incl = wave_scan(e) + run;
      const uint64_t hit = __ballot(incl > c && e > 0.0);
      const uint64_t pos = __ballot(e > 0.0);
      if (hit) {
        const uint32_t f = (uint32_t)__builtin_ctzll(hit);
        cb = base + f;
        before = __shfl(incl - e, f);
      } else if (pos) {
        const uint32_t f = 63u - (uint32_t)__builtin_clzll(pos);
        lastb = base + f;
        lastbefore = __shfl(incl - e, f);
      }
      run = __shfl(incl, 63);
    }
    if (cb == nchunks) {
      cb = lastb;
      before = lastbefore;
    }
    // within the chunk, likewise (else its last target with energy)
    const uint64_t b0 = (uint64_t)cb * L, b1 = min<uint64_t>(M, b0 + L);
    uint64_t found = b1, last = b0;
    run = before;
    for (uint64_t base = b0; base < b1 && found == b1; base += 64) {
      const double e = base + lane < b1 ? energy(y[base + lane]) : 0.0;
      const double incl = wave_scan(e) + run;
      const uint64_t hit = __ballot(incl > c && e > 0.0);
      const uint64_t pos = __ballot(e > 0.0);
      if (hit) found = base + (uint64_t)__builtin_ctzll(hit);
      else if (pos) last = base + 63u - (uint64_t)__builtin_clzll(pos);
      run = __shfl(incl, 63);
    }
    m = (uint32_t)(found < b1 ? found : last);
  }
  if (lane == 0) {
    tmap[i] = m;
    vsub[3 * (size_t)i] = v[3 * (size_t)m];
    vsub[3 * (size_t)i + 1] = v[3 * (size_t)m + 1];
    vsub[3 * (size_t)i + 2] = v[3 * (size_t)m + 2];
  }
}

uint32_t error_chunk_len(size_t M) {
  return (uint32_t)std::max<size_t>((M + kErrChunks - 1) / kErrChunks, 1);
}

uint32_t error_energy_draws(size_t M, size_t count) {
  return M <= count ? 0u : (uint32_t)(count / 2);
}

hipError_t launch_error_targets(const double2* y, size_t M, const double* v, size_t count,
                                double* E, uint32_t* tmap, double* vsub, hipStream_t stream) {
  if (!count) return hipSuccess;
  const uint32_t L = error_chunk_len(M);
  const uint32_t nchunks = (uint32_t)((M + L - 1) / L);
  k_energy_chunks<<<nchunks, kBlock, 0, stream>>>(y, (uint32_t)M, L, E);
  const unsigned blocks = (unsigned)((count + kBlock / 64 - 1) / (kBlock / 64));
  k_error_targets<<<blocks, kBlock, 0, stream>>>(y, (uint32_t)M, L, E, nchunks, v,
                                                 (uint32_t)count, error_energy_draws(M, count),
                                                 tmap, vsub);
  return hipGetLastError();
}

// One workgroup: the report from the sampled targets (weighted, or the even
// targets alone, as k_error_targets chose from the same E).
__global__ __launch_bounds__(kBlock) void k_rel_l2(const double2* __restrict__ y, uint32_t M,
                                                   const uint32_t* __restrict__ tmap,
                                                   const double2* __restrict__ ref,
                                                   uint32_t count, uint32_t ne,
                                                   const double* __restrict__ E,
                                                   uint32_t nchunks, double* __restrict__ err) {
  __shared__ double red[2][kBlock / 64];
  __shared__ double s_total;
  if (threadIdx.x < 64) {
    const double t = energy_total(E, nchunks);
    if (threadIdx.x == 0) s_total = t;
  }
  __syncthreads();
  const bool weighted = energy_mode(s_total);
  // n_u T / M, the even draws' share of the weight
  const double even = (double)(count - ne) * (s_total / (double)M);
  double num = 0.0, den = 0.0;
  for (uint32_t i = threadIdx.x; i < count; i += kBlock) {
    const double2 a = y[tmap[i]], b = ref[i];
    const double dr = a.x - b.x, di = a.y - b.y;
    if (weighted) {
      num += (dr * dr + di * di) / ((double)ne * energy(a) + even);
    } else {
      num += dr * dr + di * di;
      den += energy(b);
    }
  }
  num = wave_sum(num);
  den = wave_sum(den);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = num;
    red[1][threadIdx.x >> 6] = den;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double n = 0.0, d = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
      n += red[0][w];
      d += red[1][w];
    }
    if (weighted) err[0] = __builtin_sqrt(n);
    else err[0] = d > 0.0 ? __builtin_sqrt(n / d) : (n > 0.0 ? INFINITY : 0.0);
  }
}

hipError_t launch_rel_l2(const double2* y, size_t M, const uint32_t* tmap, const double2* ref,
                         size_t count, const double* E, double* err, hipStream_t stream) {
  const uint32_t L = error_chunk_len(M);
  k_rel_l2<<<1, kBlock, 0, stream>>>(y, (uint32_t)M, tmap, ref, (uint32_t)count,
                                     error_energy_draws(M, count), E,
                                     (uint32_t)((M + L - 1) / L), err);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------

// Zero fill with a kernel rather than hipMemsetAsync, so that device-pointer
// calls stay safe to record into a HIP graph.
__global__ __launch_bounds__(kBlock) void k_zero(double* __restrict__ x,
                                                 size_t n) {
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kBlock)
    x[i] = 0.0;
}

hipError_t launch_zero(double* x, size_t n, hipStream_t stream) {
  if (!n) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
  k_zero<<<blocks, kBlock, 0, stream>>>(x, n);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_adjoint_epilogue(
    const double2* __restrict__ b, const double* __restrict__ amp,
    const double* __restrict__ theta, uint32_t N,
    double* __restrict__ grad_amp, double* __restrict__ grad_theta) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double s, c;
  sincos(theta[n], &s, &c);
  const double2 z = b[n];
  const double pr = c * z.x + s * z.y;
  const double pi = c * z.y - s * z.x;
  grad_amp[n] = pr;
  grad_theta[n] = amp[n] * pi;
}

hipError_t launch_adjoint_epilogue(const double2* b, const double* amp,
                                   const double* theta, size_t N,
                                   double* grad_amp, double* grad_theta,
                                   hipStream_t stream) {
  if (!N) return hipSuccess;
  const unsigned blocks = (unsigned)((N + kBlock - 1) / kBlock);
  k_adjoint_epilogue<<<blocks, kBlock, 0, stream>>>(b, amp, theta, (uint32_t)N,
                                                   grad_amp, grad_theta);
  return hipGetLastError();
}

}  // namespace holo
