// Internal launch interface between the C-ABI (holo_capi.hip) and the
// kernels (holo_kernels.hip).  Every pointer here is a device pointer and
// every launch is stream-ordered.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

namespace holo {

enum WeightForm { kPolar = 0, kCartesian = 1 };

// Source set for the FP32-split kernels: sources grouped into spatial tiles
// of `tile` slots.  Slot s of tile t is geo[t*tile + s] = (dx, dy, dz, |d|^2)
// relative to the tile centre centre[t] = (cx, cy, cz, 0); perm[slot] is the
// caller's source index or -1 for a padding slot (weight 0).
struct TileSet {
  const double4* centre = nullptr;  // .w: bit-cast float2 (lateral, axial radius)
  const float4* geo = nullptr;
  const int32_t* perm = nullptr;
  const double* u = nullptr;  // the sources' FP64 positions (AoS), for exact tiles
  uint32_t n_tiles = 0;
  uint32_t tile = 0;
  bool planar = false;  // every slot has delta_z == 0
};

// Source set of the lattice kernel: the sources form a product grid
// x_i (i < nx) by y_j (j < ny) at one z, in the caller's order n = i*ny + j,
// cut into 8 x 8 blocks.  Tile t = I*(ny/8) + J holds sources (8I+a, 8J+b);
// slot 8a+b; centre[t] = (cx, cy, z, bit-cast float2 (rho_x, rho_y)) with
// rho the largest |offset| along x and y; goff[4t+g] = 2^10 (gx_2g, gx_2g+1,
// gy_2g, gy_2g+1), the FP32 offsets of rows 2g, 2g+1 and columns 2g, 2g+1
// from the centre, scaled by 2^10; perm[64t + slot] = source index.
// Bounds over all tiles (for a per-target accuracy check once per launch):
// centre x in [cx0, cx1], y in [cy0, cy1], radii <= (rho_x, rho_y), plane z.
struct GridTiles {
  const double4* centre = nullptr;
  const float4* goff = nullptr;
  const int32_t* perm = nullptr;
  const double* u = nullptr;
  uint32_t n_tiles = 0;
  double cx0 = 0, cx1 = 0, cy0 = 0, cy1 = 0, z = 0;
  float rho_x = 0, rho_y = 0;
  double px = 0, py = 0;  // pitches of a uniform lattice (16 x 16 tiles)
  // a uniform lattice's row x's (16 per row of 16 x 16 tiles) and column y's
  // (16 per column of tiles), tj tile columns, nx x ny sources (16 x 16
  // tiles; xs = nullptr: the exact path reads positions through perm)
  const double* xs = nullptr;
  const double* ys = nullptr;
  uint32_t tj = 0, nx = 0, ny = 0;
  // sources after a padded lattice (16 x 16 tiles): slots, weights, count
  const int32_t* tail_perm = nullptr;
  const float2* tail_wc = nullptr;
  uint32_t n_tail = 0;
};

struct Tuning {
  int r_f64 = 4;          // targets per lane, FP64 kernel (1, 2, 4)
  int r_f32 = 2;          // targets per lane, FP32-split kernel (1, 2, 4)
  int splits = 0;         // source splits (0 = automatic)
  int target_waves = 16384;
  bool f64_libm = false;  // FP64 kernel: device libm sqrt/div/sincos instead of hm_*
  bool f64_planar = true; // FP64 kernel: dz once per target when all sources share one z
  bool f64_rows = true;   // FP64 kernel: dx^2 once per (target, row) on such sources in rows
  bool grid = true;       // FP32-split: lattice kernel for product-grid sources
  bool grid16 = true;     // lattice kernel: 16 x 16 tiles on uniform lattices
  bool grid32 = true;     // ... as 16 x 32 tiles where the columns pair up
  bool grid64 = true;     // ... as 16 x 64 tiles where they form quads
  bool grid_series = true; // 16-row kernels: near-field pairs by the q-series, not exactly
  bool grid_near = true;   // 16-row kernels: waves off the far-field bounds take the near-field GEMM
  bool grid_rows = true;   // 16-row kernels: exact pairs read row x's and column y's from tables
  bool mc_lattice = true; // per-target FP32-split MC: table gather on product grids
  int mc_passes = 0;      // per-target MC: source ranges, one launch each (0 = automatic)
};

Tuning& tuning();

// Run-time -> compile-time dispatch: each helper calls f with a
// std::integral_constant carrying the value.
template <int V>
using IC = std::integral_constant<int, V>;

template <typename F>
inline void with_bool(bool b, F&& f) {
  if (b) f(IC<1>{});
  else f(IC<0>{});
}

template <typename F>
inline void with_sign(int sign, F&& f) {
  if (sign > 0) f(IC<+1>{});
  else f(IC<-1>{});
}

template <typename F>
inline void with_r(int R, F&& f) {
  switch (R) {
    case 1: f(IC<1>{}); break;
    case 4: f(IC<4>{}); break;
    default: f(IC<2>{}); break;
  }
}


// Number of source splits S chosen for a launch over M targets with R
// targets per lane and `units` source units (sources or tiles), at most
// max_splits.
int choose_splits(size_t M, int R, size_t units, size_t min_units_per_split,
                  size_t max_splits = 64);

// Source splits of a launch over a few hundred targets (the error report's
// subset): enough workgroups to fill the 256 CUs.
constexpr size_t kFewTargetSplits = 1024;

// y (interleaved complex, M) = sum_n pair(n, m).  Polar weights: w0 = amp,
// w1 = phase.  Cartesian: w0 = interleaved complex weights, w1 unused.
// obliquity: every pair's amplitude is also multiplied by |dz| / d.
// uniform_z: every source has the z of source 0 (same results, one operation
// fewer per pair).  `partial` holds partial_cap double2; S source splits need
// S*M of them.
// row > 0: the sources form rows of `row` consecutive sources with one x
// (used with uniform_z).  few_targets: one target per lane and up to
// kFewTargetSplits source splits, for M of a few hundred.
hipError_t launch_superpose_f64(int sign, WeightForm wf, bool obliquity, bool uniform_z,
                                size_t row, const double* w0, const double* w1, const double* u,
                                size_t N,
                                const double* v, size_t M, double K,
                                double2* y, double2* partial,
                                size_t partial_cap, hipStream_t stream,
                                bool few_targets = false);

// The same sum for sources whose layout the host has not scanned (device
// pointers): launch_source_layout writes it to `layout` (4 words of device
// memory) and the kernel reads it there, taking the uniform-z and rows paths
// of launch_superpose_f64 where they apply (the same bits).
hipError_t launch_superpose_f64_auto(int sign, WeightForm wf, bool obliquity,
                                     const double* w0, const double* w1, const double* u,
                                     size_t N, const double* v, size_t M, double K, double2* y,
                                     double2* partial, size_t partial_cap, uint32_t* layout,
                                     hipStream_t stream);
// layout[0] bit 0: every source has the z of source 0; bit 1: rows of
// layout[1] (>= 4) consecutive sources that share one x, the last row
// possibly shorter; the same rules as the host scan.  Bits not in `allow`
// stay clear.
hipError_t launch_source_layout(const double* u, size_t N, uint32_t* layout,
                                hipStream_t stream, unsigned allow = 3u);

// y[m] = sum_s partial[s*M + m] in fixed order (nothing to do for S <= 1).
hipError_t launch_reduce_partials(int S, size_t M, const double2* partial, double2* y,
                                  hipStream_t stream);

// FP32-split: weights wc (float2 per slot, polar: launch_tile_weights with
// polar_out), tiles as above.
hipError_t launch_superpose_f32s(int sign, bool obliquity, const TileSet& ts,
                                 const float2* wc, const double* v, size_t M,
                                 double lambda, double2* y, double2* partial,
                                 size_t partial_cap, hipStream_t stream);

// Lattice kernel (FP32-split on product-grid tiles).  wc: per-slot FP32
// weights (launch_tile_weights with the grid perm).  launch_grid_weights
// lays them out as the MFMA operands, divided by tscale[t] (a power of
// two): wh and wl, 8 x f16 per tile and lane each (the f16 hi and lo parts
// of the weights and the cross-term weights i gx W).
hipError_t launch_grid_weights(const float2* wc, const float4* goff, uint32_t n_tiles,
                               uint4* wh, uint4* wl, float* tscale, hipStream_t stream);
hipError_t launch_superpose_grid(int sign, bool obliquity, const GridTiles& gt,
                                 const float2* wc, const uint4* wh, const uint4* wl,
                                 const float* tscale, const double* v, size_t M,
                                 double lambda, double2* y, double2* partial,
                                 size_t partial_cap, hipStream_t stream);
// 16 x 16 tiles of a uniform lattice (pitches gt.px, gt.py; gt.goff unused):
// wc holds 256 slots per tile (slot 16a + b), launch_grid16_weights lays them
// out as wq (6 x 8 f16 per tile and lane), wqn (the near-field GEMM's 4 x 8 f16
// per tile and lane; nullable) and tscale; tiles group T .. group T + group - 1
// share one tscale (group 1, 2 or 4).  launch_superpose_grid16 with nkc = 1
// takes the 16 x 16 tiles gt (pairs and halves unused); with nkc = 2 it takes
// the 16 x 32 tiles formed by pairs of 16 x 16 tiles (gt.n_tiles of them, their
// centres in gt.centre, rho_y and the bounds those of the pairs), reading the
// same wc, wq, perm and tscale (from launch_grid16_weights with group 2), and
// halves describes the same sources as 16 x 16 tiles (2 gt.n_tiles of them): a
// wave whose targets fail the 16 x 32 far-field bounds takes those.  With
// nkc = 4, gt holds the 16 x 64 tiles formed by quads (group 4), pairs the
// same sources as 16 x 32 tiles (2 gt.n_tiles) and halves as 16 x 16 tiles
// (4 gt.n_tiles): a wave failing the 16 x 64 bounds takes the 16 x 32 tiles if
// it passes theirs, else the 16 x 16 ones.
hipError_t launch_grid16_weights(const float2* wc, uint32_t n_tiles, int group, uint4* wq,
                                 uint4* wqn, float* tscale, hipStream_t stream);
// wqn null: no near-field GEMM (the waves off the far-field bounds take the
// series and exact paths per (target, tile))
hipError_t launch_superpose_grid16(int sign, bool obliquity, int nkc, const GridTiles& gt,
                                   const GridTiles& pairs, const GridTiles& halves,
                                   const float2* wc, const uint4* wq, const uint4* wqn,
                                   const float* tscale, const double* v, size_t M,
                                   double lambda, double2* y, double2* partial,
                                   size_t partial_cap, hipStream_t stream);
// Source splits the lattice kernel uses for M targets over n_tiles tiles.
int grid_splits(size_t M, size_t n_tiles);

// wc[slot] = a e^{i phi} (polar input) or w (cartesian input) of source
// perm[slot] as (re, im); with polar_out as (|w|, arg w / 2 pi in [-1/2, 1/2)).
hipError_t launch_tile_weights(WeightForm wf, const double* w0,
                               const double* w1, const int32_t* perm,
                               size_t n_slots, float2* wc, hipStream_t stream,
                               bool polar_out = false);

// Monte Carlo: y[m] = (N/K) sum_k pair(n_{m,k}, m).
// FP64: pos = (x,y,z,amp) per source, ph = phase per source.
// FP32-split: pos = (x,y,z, bitcast float2 weight).
constexpr unsigned kMcGroup = 1024;  // targets sharing one subset (batched MC)

// Sources on an x-slow product grid (x depends on n / ny only, y on n % ny
// only, one z), found on the device by launch_mc_lattice_probe:
//   info[0] = ny (0: not found), info[1] = the number of leading complete
//   rows (sources n < info[1] * ny lie on the grid; the rest are read from
//   the packed records);
//   xy[j] = y of column j (j < ny <= kLatticeMaxRow), xy[kLatticeMaxRow] = z,
//   xy[kLatticeMaxRow + 1 + i] = x of row i;  wc[n] = the FP32 weight.
// The per-target FP32-split kernel then gathers 8 bytes per sample instead
// of a 32-byte record, with bit-identical results.
constexpr uint32_t kLatticeMaxRow = 65536;
struct McLattice {
  const uint32_t* info = nullptr;
  const float2* wc = nullptr;
  const double* xy = nullptr;
};

// Per-target Monte Carlo runs as one launch per source range of about
// kMcPassBytes of gathered data, which one XCD's 4 MB L2 holds beside the
// streamed targets (on the 1024^2 lattice, K = 4096: 39.1 ms in one launch,
// 23.0 / 22.9 / 25.9 / 32.9 ms in 2 / 3 / 4 / 6 ranges): the number of ranges.
constexpr size_t kMcPassBytes = size_t(3) << 20;
uint32_t mc_passes(size_t N, bool fp64);

hipError_t launch_mc(int sign, bool fp64, bool shared, bool obliquity, const double4* pos,
                     const double* ph, size_t N, const double* v, size_t M,
                     uint32_t K, uint64_t seed, double K_wave, double lambda,
                     const McLattice& lat, double2* y, hipStream_t stream);

// The FP64 estimator at `count` targets vsub whose indices (Philox counters)
// are tmap[i]: per target, or (shared) of the target's group; pos/ph packed
// for FP64 by launch_mc_pack.  For the error report of an FP32-split call.
hipError_t launch_mc_subset(int sign, bool shared, bool obliquity, const double4* pos,
                            const double* ph, size_t N, const double* vsub, size_t count,
                            const uint32_t* tmap, uint32_t K, uint64_t seed, double K_wave,
                            double lambda, double2* y, double2* partial, size_t partial_cap,
                            hipStream_t stream);

// wc may be NULL; otherwise it also receives the FP32 weights (FP32-split).
hipError_t launch_mc_pack(bool fp64, const double* amp, const double* phase,
                          const double* u, size_t N, double4* pos, double* ph, float2* wc,
                          hipStream_t stream);

// info: 2 words, xy: kLatticeMaxRow + 1 + N doubles (device memory).
hipError_t launch_mc_lattice_probe(const double* u, size_t N, uint32_t* info, double* xy,
                                   hipStream_t stream);

// Per-target indices of targets m0..m0+count-1, or (shared) the subsets of
// groups m0..m0+count-1.
hipError_t launch_mc_indices(uint64_t seed, uint64_t N, size_t m0,
                             size_t count, uint32_t K, bool shared, int64_t* idx,
                             hipStream_t stream);

// max |y| into *max_out (device), then amp/phase = polar(y / max) (or y when
// max < 1e-6).  scratch: >= 1024 doubles.
hipError_t launch_normalize_polar(const double2* y, size_t M, double* amp,
                                  double* phase, double* max_out,
                                  double* scratch, hipStream_t stream);

hipError_t launch_zero(double* x, size_t n, hipStream_t stream);

// Error report of an FP32-split result y (M values): launch_error_targets
// picks count <= kErrTargets targets, half drawn by the energy of y and half
// spread evenly (every target when M <= count) (tmap[i], positions vsub[i];
// E: kErrChunks doubles of partial energies, kept for launch_rel_l2), the
// caller evaluates the FP64 sum ref at vsub, and launch_rel_l2 writes the
// report to *err (device): the importance-weighted estimate of the whole
// result's relative L2 error, or with no finite positive energy the even
// targets' relative L2.  Fixed summation orders.
constexpr size_t kErrTargets = 256;
constexpr size_t kErrChunks = 256;
hipError_t launch_error_targets(const double2* y, size_t M, const double* v, size_t count,
                                double* E, uint32_t* tmap, double* vsub, hipStream_t stream);
hipError_t launch_rel_l2(const double2* y, size_t M, const uint32_t* tmap, const double2* ref,
                         size_t count, const double* E, double* err, hipStream_t stream);

// Adjoint epilogue: grad_amp = Re(e^{-i theta} b), grad_theta = A Im(...).
hipError_t launch_adjoint_epilogue(const double2* b, const double* amp,
                                   const double* theta, size_t N,
                                   double* grad_amp, double* grad_theta,
                                   hipStream_t stream);

}  // namespace holo
