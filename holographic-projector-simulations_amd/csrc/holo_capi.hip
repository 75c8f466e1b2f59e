// C ABI of libholo.so (declared in include/holo.h).
//
// Host-side orchestration only: argument checks, device buffers owned by the
// context, host<->device staging, the FP32-split tiling of source positions,
// and launches of the kernels in holo_kernels.hip.  Every failure returns a
// status and records a message; nothing exits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/holo.h"
#include "holo_internal.hpp"
#include "holo_math.hpp"

namespace {

thread_local std::string g_error;

holo_status fail(holo_status s, const std::string& msg) {
  g_error = msg;
  return s;
}

#define HOLO_HIP(expr)                                                     \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess)                                                  \
      return fail(e_ == hipErrorOutOfMemory ? HOLO_ENOMEM : HOLO_EDEVICE,  \
                  std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define HOLO_TRY(expr)                 \
  do {                                 \
    holo_status s_ = (expr);           \
    if (s_ != HOLO_OK) return s_;      \
  } while (0)

constexpr size_t kMaxCount = (size_t(1) << 31) - 1;

// Growable device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int device = 0;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) {
      int cur = 0;
      (void)hipGetDevice(&cur);
      (void)hipSetDevice(device);
      (void)hipFree(ptr);
      (void)hipSetDevice(cur);
    }
    ptr = nullptr;
    bytes = 0;
  }
};

holo_status ensure(DevBuf& b, size_t bytes, int device) {
  if (bytes <= b.bytes && b.ptr) return HOLO_OK;
  b.release();
  b.device = device;
  const size_t want = std::max<size_t>(bytes, 256);
  HOLO_HIP(hipMalloc(&b.ptr, want));
  b.bytes = want;
  return HOLO_OK;
}

}  // namespace

struct holo_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // staging buffers for host-pointer calls and scratch
  DevBuf amp, phase, u, v, y, w, partial, small, b, pos, ph;
  DevBuf mc_wc, mc_xy, mc_info;  // product-grid Monte Carlo (McLattice)
  DevBuf layout;                 // source layout of device-pointer FP64 calls
  DevBuf errb;                   // error report: subset, FP64 values, partials
  DevBuf pos64, ph64;            // FP64 Monte Carlo records for the error report
  // recorded on the previous stream when holo_set_stream switches streams
  hipEvent_t switched = nullptr;
};

struct holo_plan {
  int device = 0;
  size_t N = 0;
  DevBuf u;  // AoS xyz, 3N doubles
  bool uniform_z = false;  // every source has the z of source 0
  size_t row = 0;          // sources in rows of `row` with one x (0: not)
  // FP32-split tiling
  unsigned tile = 16;
  size_t n_tiles = 0;
  double max_offset = 0.0;
  bool planar = true;
  // The spatial tiling for k_superpose_f32s.  A lattice plan defers it to
  // the first call that needs it (the lattice kernel does not).
  mutable bool tiles_built = false;
  mutable std::mutex tiles_mu;
  mutable DevBuf centre, geo, perm, wc;
  // lattice tiling (product-grid sources, 8 x 8 tiles) for the MFMA kernel
  size_t grid_nx = 0, grid_ny = 0, grid_tiles = 0;
  double grid_cx0 = 0, grid_cx1 = 0, grid_cy0 = 0, grid_cy1 = 0, grid_z = 0;
  float grid_rho_x = 0, grid_rho_y = 0;
  DevBuf g_centre, g_goff, g_perm, g_wc, g_wh, g_wl, g_ts;
  // 16 x 16 tiles of the same lattice when its pitches are uniform, over the
  // lattice padded with weight-0 slots to multiples of 16 (g16_tiles = 0
  // otherwise)
  size_t g16_tiles = 0;
  // sources after a padded lattice, summed exactly: their slots follow the
  // tiles' in g16_perm and g16_wc
  size_t g16_tail = 0;
  double g16_px = 0, g16_py = 0;
  double g16_cx0 = 0, g16_cx1 = 0, g16_cy0 = 0, g16_cy1 = 0;
  float g16_rho_x = 0, g16_rho_y = 0;
  DevBuf g16_centre, g16_perm, g16_wc, g16_wq, g16_wqn, g16_ts;
  // their row x's and column y's (16 TI + 16 TJ) for the exact path
  size_t g16_tj = 0, g16_nx = 0, g16_ny = 0;
  DevBuf g16_axes;
  // the 16 x 32 tiles formed by pairs of 16 x 16 tiles (16x16 tiles 2T, 2T+1)
  // when the padded count of tile columns is even (g32_tiles = 0 otherwise)
  size_t g32_tiles = 0;
  double g32_cx0 = 0, g32_cx1 = 0, g32_cy0 = 0, g32_cy1 = 0;
  float g32_rho_y = 0;
  DevBuf g32_centre;
  // the 16 x 64 tiles formed by quads of 16 x 16 tiles (4T .. 4T + 3) when the
  // padded count of tile columns is a multiple of 4 (g64_tiles = 0 otherwise)
  size_t g64_tiles = 0;
  double g64_cx0 = 0, g64_cx1 = 0, g64_cy0 = 0, g64_cy1 = 0;
  float g64_rho_y = 0;
  DevBuf g64_centre;
  // Every call rewrites the weight buffers above.  `ran` is recorded after
  // each call on its stream; a call on another stream waits for it first.
  mutable hipEvent_t ran = nullptr;
  mutable hipStream_t ran_on = nullptr;
  ~holo_plan() {
    if (ran) {
      int cur = 0;
      (void)hipGetDevice(&cur);
      (void)hipSetDevice(device);
      (void)hipEventDestroy(ran);
      (void)hipSetDevice(cur);
    }
  }
};

namespace {

// ---- scratch of holo_transform_full_noalloc --------------------------------
// The split partials and source-layout words of the no-allocation entry point
// live in this file-scope device array: one copy per device, allocated when
// the code object loads on that device and held until the process ends.  In
// host code `scratch` names a host-side stand-in, not device memory: kernels
// and copies get the address hipGetSymbolAddress returns for the current
// device, kept per device in g_scratch.
constexpr size_t kNoallocScratchBytes = HOLO_NOALLOC_SCRATCH_BYTES;
constexpr size_t kNoallocLayoutBytes = 256;  // 4 layout words, padded
constexpr int kMaxDevices = 64;
static_assert(kNoallocScratchBytes % 256 == 0 && kNoallocScratchBytes > kNoallocLayoutBytes,
              "scratch size");
__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current.  The
// first call on a device looks it up; no call allocates.
holo_status noalloc_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(HOLO_EDEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    HOLO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(HOLO_EDEVICE, "hipGetSymbolAddress returned no address for scratch");
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return HOLO_OK;
}

holo_status use_device(holo_ctx* ctx) {
  if (!ctx) return fail(HOLO_EINVAL, "null context");
  HOLO_HIP(hipSetDevice(ctx->device));
  return HOLO_OK;
}

holo_status check_common(holo_ctx* ctx, int precision, int sign, double lambda,
                         size_t N, size_t M) {
  if (!ctx) return fail(HOLO_EINVAL, "null context");
  if (precision != HOLO_FP64 && precision != HOLO_FP32_SPLIT)
    return fail(HOLO_EINVAL, "precision must be HOLO_FP64 or HOLO_FP32_SPLIT");
  if (sign != 1 && sign != -1) return fail(HOLO_EINVAL, "sign must be +1 or -1");
  if (!(lambda > 0.0) || !std::isfinite(lambda))
    return fail(HOLO_EINVAL, "lambda must be positive and finite");
  if (N > kMaxCount || M > kMaxCount)
    return fail(HOLO_EINVAL, "N and M must be below 2^31");
  return HOLO_OK;
}

holo_status check_flags(unsigned flags, unsigned allowed) {
  if (flags & ~allowed) return fail(HOLO_EINVAL, "unsupported flag bits for this call");
  return HOLO_OK;
}

// Device view of an input array: the caller's pointer in device mode, a
// staged copy otherwise.
template <typename T>
holo_status stage_in(holo_ctx* ctx, DevBuf& buf, const T* src, size_t count,
                     bool dev, const T** out) {
  if (count == 0) {
    *out = nullptr;
    return HOLO_OK;
  }
  if (!src) return fail(HOLO_EINVAL, "null input array");
  if (dev) {
    *out = src;
    return HOLO_OK;
  }
  HOLO_TRY(ensure(buf, count * sizeof(T), ctx->device));
  HOLO_HIP(hipMemcpyAsync(buf.ptr, src, count * sizeof(T), hipMemcpyHostToDevice,
                          ctx->stream));
  *out = static_cast<const T*>(buf.ptr);
  return HOLO_OK;
}

template <typename T>
holo_status out_buf(holo_ctx* ctx, DevBuf& buf, T* dst, size_t count, bool dev,
                    T** out) {
  if (count && !dst) return fail(HOLO_EINVAL, "null output array");
  if (dev) {
    *out = dst;
    return HOLO_OK;
  }
  HOLO_TRY(ensure(buf, std::max<size_t>(count, 1) * sizeof(T), ctx->device));
  *out = static_cast<T*>(buf.ptr);
  return HOLO_OK;
}

template <typename T>
holo_status finish_out(holo_ctx* ctx, T* host, const T* devp, size_t count,
                       bool dev) {
  if (!dev && count)
    HOLO_HIP(hipMemcpyAsync(host, devp, count * sizeof(T), hipMemcpyDeviceToHost,
                            ctx->stream));
  if (!dev) HOLO_HIP(hipStreamSynchronize(ctx->stream));
  return HOLO_OK;
}

// ---------------------------------------------------------------------------
// Error report of a result r at M targets (holo.h): the FP64 formulation at
// count = min(M, kErrTargets) targets drawn by the energy of r, then the
// weighted mean square of their relative errors on the device.  The scratch
// is one context buffer.
struct ErrScratch {
  double* err = nullptr;      // the report, when the caller's pointer is host memory
  double* energy = nullptr;   // kErrChunks partial sums of |r|^2
  uint32_t* tmap = nullptr;   // sampled targets m_i
  double* vsub = nullptr;     // target positions at m_i
  double2* ref = nullptr;     // FP64 values at m_i
  double2* partial = nullptr; // split partials of the FP64 launch
  size_t partial_cap = 0;
};

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

holo_status err_scratch(holo_ctx* ctx, ErrScratch* e) {
  const size_t n = holo::kErrTargets, np = holo::kFewTargetSplits * n;
  const size_t o_energy = 256, o_tmap = o_energy + align256(holo::kErrChunks * sizeof(double)),
               o_vsub = o_tmap + align256(n * sizeof(uint32_t)),
               o_ref = o_vsub + align256(3 * n * sizeof(double)),
               o_part = o_ref + align256(n * sizeof(double2)),
               total = o_part + np * sizeof(double2);
  HOLO_TRY(ensure(ctx->errb, total, ctx->device));
  char* b = static_cast<char*>(ctx->errb.ptr);
  e->err = reinterpret_cast<double*>(b);
  e->energy = reinterpret_cast<double*>(b + o_energy);
  e->tmap = reinterpret_cast<uint32_t*>(b + o_tmap);
  e->vsub = reinterpret_cast<double*>(b + o_vsub);
  e->ref = reinterpret_cast<double2*>(b + o_ref);
  e->partial = reinterpret_cast<double2*>(b + o_part);
  e->partial_cap = np;
  return HOLO_OK;
}

// Writes the report of result y (device, M values at targets v) to err_out
// (host or device pointer as `dev`; nothing when NULL): 0 when fp64, else
// run(e, count) fills e.ref with the FP64 values at e.vsub (count of them)
// first.  A host err_out is written by a queued copy: the caller waits.
holo_status report_zero(holo_ctx* ctx, double* err_out, bool dev) {
  if (!err_out) return HOLO_OK;
  if (dev) HOLO_HIP(holo::launch_zero(err_out, 1, ctx->stream));
  else *err_out = 0.0;
  return HOLO_OK;
}

template <typename F>
holo_status report_error(holo_ctx* ctx, bool fp64, const double* v, size_t M, const double* y,
                         double* err_out, bool dev, F&& run) {
  if (!err_out) return HOLO_OK;
  if (fp64 || M == 0) return report_zero(ctx, err_out, dev);
  ErrScratch e;
  HOLO_TRY(err_scratch(ctx, &e));
  const size_t count = std::min(M, holo::kErrTargets);
  const double2* r = reinterpret_cast<const double2*>(y);
  HOLO_HIP(holo::launch_error_targets(r, M, v, count, e.energy, e.tmap, e.vsub, ctx->stream));
  HOLO_TRY(run(e, count));
  HOLO_HIP(holo::launch_rel_l2(r, M, e.tmap, e.ref, count, e.energy, dev ? err_out : e.err,
                               ctx->stream));
  if (!dev)
    HOLO_HIP(hipMemcpyAsync(err_out, e.err, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  return HOLO_OK;
}

size_t split_elems_f64(size_t M, size_t N) {
  const int R = holo::tuning().r_f64;
  const int S = N ? holo::choose_splits(M, R == 1 || R == 4 ? R : 2, N, 1024) : 1;
  return S > 1 ? (size_t)S * M : 0;
}

size_t split_elems_f32(size_t M, size_t n_tiles) {
  const int R = holo::tuning().r_f32;
  const int S = n_tiles ? holo::choose_splits(M, R == 1 || R == 4 ? R : 2, n_tiles, 64) : 1;
  return S > 1 ? (size_t)S * M : 0;
}

// ---------------------------------------------------------------------------
// Spatial tiling of source positions for the FP32-split kernels.
//
// Sources are binned into cells of a regular grid over their bounding box,
// sized so that a cell holds about `tile` sources (axes whose extent would
// get fewer than ~1.5 cells are dropped, so slabs and lines are binned in 2-D
// or 1-D).  Cells are visited in Morton order; each cell's sources, in input
// order, fill whole tiles, and the partial remainders of consecutive cells
// are merged while the merged tile stays within two cell widths.  The last
// tile is padded with weight-0 slots.  The tile centre is the midpoint of
// its sources' bounding box.  A regular n x n lattice with n divisible by
// sqrt(tile) gets exact sqrt(tile) x sqrt(tile) blocks and no padding.
struct HostTiles {
  std::vector<double4> centre;
  std::vector<float4> geo;
  std::vector<int32_t> perm;
  double max_offset = 0.0;
  bool planar = true;
};

uint64_t spread_bits21(uint64_t x) {
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

void emit_tile(const double* u, const int32_t* idx, size_t cnt, unsigned T,
               HostTiles& out) {
  double blo[3] = {INFINITY, INFINITY, INFINITY};
  double bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t s = 0; s < cnt; ++s)
    for (int d = 0; d < 3; ++d) {
      blo[d] = std::min(blo[d], u[3 * (size_t)idx[s] + d]);
      bhi[d] = std::max(bhi[d], u[3 * (size_t)idx[s] + d]);
    }
  double c[3];
  for (int d = 0; d < 3; ++d) c[d] = 0.5 * (blo[d] + bhi[d]);
  double rho_xy = 0.0, rho_z = 0.0;
  for (size_t s = 0; s < cnt; ++s) {
    const double* p = &u[3 * (size_t)idx[s]];
    rho_xy = std::max(rho_xy, std::hypot(p[0] - c[0], p[1] - c[1]));
    rho_z = std::max(rho_z, std::fabs(p[2] - c[2]));
  }
  // radii rounded up to floats, packed into .w
  const float2 rho{std::nextafter((float)rho_xy, INFINITY), std::nextafter((float)rho_z, INFINITY)};
  out.centre.push_back(double4{c[0], c[1], c[2], __builtin_bit_cast(double, rho)});
  for (size_t s = 0; s < T; ++s) {
    if (s < cnt) {
      const int32_t n = idx[s];
      const float fx = (float)(u[3 * (size_t)n] - c[0]);
      const float fy = (float)(u[3 * (size_t)n + 1] - c[1]);
      const float fz = (float)(u[3 * (size_t)n + 2] - c[2]);
      const double dd = (double)fx * fx + (double)fy * fy + (double)fz * fz;
      out.geo.push_back(float4{fx, fy, fz, (float)dd});
      out.perm.push_back(n);
      if (fz != 0.f) out.planar = false;
      out.max_offset = std::max(out.max_offset, std::sqrt(dd));
    } else {
      out.geo.push_back(float4{0.f, 0.f, 0.f, 0.f});
      out.perm.push_back(-1);
    }
  }
}

void build_tiles(const double* u, size_t N, unsigned T, HostTiles& out) {
  out = HostTiles();
  if (N == 0) return;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t n = 0; n < N; ++n)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], u[3 * n + d]);
      hi[d] = std::max(hi[d], u[3 * n + d]);
    }
  double ext[3], emax = 0.0;
  for (int d = 0; d < 3; ++d) {
    ext[d] = hi[d] - lo[d];
    emax = std::max(emax, ext[d]);
  }
  bool act[3];
  for (int d = 0; d < 3; ++d) act[d] = ext[d] > 0.0 && ext[d] > 1e-9 * emax;
  size_t cells[3] = {1, 1, 1};
  double h = 0.0;
  if (N > T) {
    for (int iter = 0; iter < 3; ++iter) {
      int na = 0;
      double vol = 1.0;
      for (int d = 0; d < 3; ++d)
        if (act[d]) {
          ++na;
          vol *= ext[d];
        }
      if (!na) break;
      h = std::pow(vol * (double)T / (double)N, 1.0 / na);
      bool changed = false;
      for (int d = 0; d < 3; ++d)
        if (act[d] && ext[d] / h < 1.5) {
          act[d] = false;
          changed = true;
        }
      if (!changed) break;
    }
    for (int d = 0; d < 3; ++d)
      if (act[d]) cells[d] = (size_t)std::max(1.0, std::min(std::round(ext[d] / h), 2097151.0));
  }
  std::vector<uint64_t> key(N);
  for (size_t n = 0; n < N; ++n) {
    uint64_t k = 0;
    for (int d = 0; d < 3; ++d) {
      uint64_t c = 0;
      if (cells[d] > 1) {
        const double f = (u[3 * n + d] - lo[d]) / ext[d] * (double)cells[d];
        c = (uint64_t)std::max(0.0, std::floor(f));
        if (c >= cells[d]) c = cells[d] - 1;
      }
      k |= spread_bits21(c) << d;
    }
    key[n] = k;
  }
  std::vector<int32_t> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  const double merge_limit = 2.0 * h;
  std::vector<int32_t> pending;
  double plo[3], phi[3];
  size_t i = 0;
  while (i < N) {
    size_t j = i;
    while (j < N && key[order[j]] == key[order[i]]) ++j;  // one cell: [i, j)
    const size_t full = (j - i) / T * T;
    for (size_t t0 = i; t0 < i + full; t0 += T) emit_tile(u, &order[t0], T, T, out);
    if (full < j - i) {
      double llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (size_t s = i + full; s < j; ++s)
        for (int d = 0; d < 3; ++d) {
          llo[d] = std::min(llo[d], u[3 * (size_t)order[s] + d]);
          lhi[d] = std::max(lhi[d], u[3 * (size_t)order[s] + d]);
        }
      bool merge = !pending.empty() && pending.size() + (j - i - full) <= T;
      for (int d = 0; d < 3 && merge; ++d)
        merge = std::max(phi[d], lhi[d]) - std::min(plo[d], llo[d]) <= merge_limit;
      if (!merge) {
        if (!pending.empty()) emit_tile(u, pending.data(), pending.size(), T, out);
        pending.clear();
        for (int d = 0; d < 3; ++d) {
          plo[d] = llo[d];
          phi[d] = lhi[d];
        }
      } else {
        for (int d = 0; d < 3; ++d) {
          plo[d] = std::min(plo[d], llo[d]);
          phi[d] = std::max(phi[d], lhi[d]);
        }
      }
      pending.insert(pending.end(), order.begin() + (i + full), order.begin() + j);
    }
    i = j;
  }
  if (!pending.empty()) emit_tile(u, pending.data(), pending.size(), T, out);
}

// Product-grid detection for the lattice kernel: positions n = i*ny + j with
// x depending on the row i only, y on the column j only, and one z (the
// reference's plane generators, cuda/init.h:89-158, x slow as in
// cuda/macros.h:59).  nx and ny must be multiples of 8; the grid is cut into
// 8 x 8 tiles t = I*(ny/8) + J, slot 8a + b = source (8I+a, 8J+b).  When the
// positions are x_0 + i p_x, y_0 + j p_y (to 1e-9 of a pitch), it is also cut
// into 16 x 16 tiles t = I*(TJ) + J, TJ = ceil(ny/16), slot 16a + b = source
// (16I+a, 16J+b) or padding (-1) beyond the lattice, centred at the lattice
// points x_0 + (16I + 7.5) p_x, y_0 + (16J + 7.5) p_y (with padding, an odd
// TJ of 5 or more takes one more column of tiles).  When TJ is even,
// tiles 2T and 2T + 1 (the same rows, columns 32J' .. 32J' + 31) also form
// the 16 x 32 tile T, centred at x_0 + (16I + 7.5) p_x, y_0 + (32J' + 15.5) p_y.
struct HostGrid {
  size_t nx = 0, ny = 0;
  double cx0 = INFINITY, cx1 = -INFINITY, cy0 = INFINITY, cy1 = -INFINITY, z = 0;
  float rho_x = 0, rho_y = 0;
  std::vector<double4> centre;
  std::vector<float4> goff;
  std::vector<int32_t> perm;
  double px = 0, py = 0;  // nonzero: 16 x 16 tiles below
  double cx0_16 = INFINITY, cx1_16 = -INFINITY, cy0_16 = INFINITY, cy1_16 = -INFINITY;
  float rho16_x = 0, rho16_y = 0;
  std::vector<double4> centre16;
  std::vector<int32_t> perm16;
  // the 16 x 16 tiles' row x's (16 TI, row a: x of source a ny) followed by
  // their column y's (16 TJ, column b: y of source b), extended over the
  // padding by the lattice's pitches; tj16 = TJ
  std::vector<double> axes16;
  size_t tj16 = 0;
  double cx0_32 = INFINITY, cx1_32 = -INFINITY, cy0_32 = INFINITY, cy1_32 = -INFINITY;
  float rho32_y = 0;
  std::vector<double4> centre32;
  double cx0_64 = INFINITY, cx1_64 = -INFINITY, cy0_64 = INFINITY, cy1_64 = -INFINITY;
  float rho64_y = 0;
  std::vector<double4> centre64;
  std::vector<int32_t> tail;  // sources after the lattice (padded lattices)
};

// Sides that are not multiples of 16 are padded to the next multiple with
// weight-0 slots (perm -1) at the lattice points beyond the edge.
void build_grid16(const double* u, HostGrid& g) {
  const size_t nx = g.nx, ny = g.ny;
  const double x0 = u[0], y0 = u[1];
  const double px = (u[3 * (nx - 1) * ny] - x0) / (double)(nx - 1);
  const double py = (u[3 * (ny - 1) + 1] - y0) / (double)(ny - 1);
  if (!(px != 0.0 && py != 0.0 && std::isfinite(px) && std::isfinite(py))) return;
  for (size_t i = 0; i < nx; ++i)
    if (!(std::fabs(u[3 * i * ny] - (x0 + (double)i * px)) <= 1e-9 * std::fabs(px))) return;
  for (size_t j = 0; j < ny; ++j)
    if (!(std::fabs(u[3 * j + 1] - (y0 + (double)j * py)) <= 1e-9 * std::fabs(py))) return;
  g.px = px;
  g.py = py;
  g.rho16_x = std::nextafter((float)(7.5 * std::fabs(px)), INFINITY);
  g.rho16_y = std::nextafter((float)(7.5 * std::fabs(py)), INFINITY);
  const float2 rho{g.rho16_x, g.rho16_y};
  // an odd number of tile columns (5 or more) takes one more, of padding only,
  // so that they pair into 16 x 32 tiles (1000 columns: 64 instead of 63)
  const size_t TI = (nx + 15) / 16, TJ = (ny + 15) / 16 + ((ny + 15) / 16 % 2 && ny > 64);
  g.centre16.reserve(TI * TJ);
  g.perm16.reserve(256 * TI * TJ);
  g.tj16 = TJ;
  g.axes16.resize(16 * (TI + TJ));
  for (size_t a = 0; a < 16 * TI; ++a)
    g.axes16[a] = a < nx ? u[3 * a * ny] : x0 + (double)a * px;
  for (size_t b = 0; b < 16 * TJ; ++b)
    g.axes16[16 * TI + b] = b < ny ? u[3 * b + 1] : y0 + (double)b * py;
  for (size_t I = 0; I < TI; ++I)
    for (size_t J = 0; J < TJ; ++J) {
      const double cx = x0 + ((double)(16 * I) + 7.5) * px;
      const double cy = y0 + ((double)(16 * J) + 7.5) * py;
      g.centre16.push_back(double4{cx, cy, g.z, __builtin_bit_cast(double, rho)});
      g.cx0_16 = std::min(g.cx0_16, cx);
      g.cx1_16 = std::max(g.cx1_16, cx);
      g.cy0_16 = std::min(g.cy0_16, cy);
      g.cy1_16 = std::max(g.cy1_16, cy);
      for (size_t a = 16 * I; a < 16 * I + 16; ++a)
        for (size_t b = 16 * J; b < 16 * J + 16; ++b)
          g.perm16.push_back(a < nx && b < ny ? (int32_t)(a * ny + b) : -1);
    }
  if (TJ % 2) return;
  g.rho32_y = std::nextafter((float)(15.5 * std::fabs(py)), INFINITY);
  const float2 rho32{g.rho16_x, g.rho32_y};
  g.centre32.reserve(TI * TJ / 2);
  for (size_t I = 0; I < TI; ++I)
    for (size_t J = 0; J < TJ / 2; ++J) {
      const double cx = x0 + ((double)(16 * I) + 7.5) * px;
      const double cy = y0 + ((double)(32 * J) + 15.5) * py;
      g.centre32.push_back(double4{cx, cy, g.z, __builtin_bit_cast(double, rho32)});
      g.cx0_32 = std::min(g.cx0_32, cx);
      g.cx1_32 = std::max(g.cx1_32, cx);
      g.cy0_32 = std::min(g.cy0_32, cy);
      g.cy1_32 = std::max(g.cy1_32, cy);
    }
  // quads of tiles 4T .. 4T + 3: 16 x 64 tiles centred at
  // x_0 + (16I + 7.5) p_x, y_0 + (64J' + 31.5) p_y (no extra padding for them)
  if (TJ % 4) return;
  g.rho64_y = std::nextafter((float)(31.5 * std::fabs(py)), INFINITY);
  const float2 rho64{g.rho16_x, g.rho64_y};
  g.centre64.reserve(TI * TJ / 4);
  for (size_t I = 0; I < TI; ++I)
    for (size_t J = 0; J < TJ / 4; ++J) {
      const double cx = x0 + ((double)(16 * I) + 7.5) * px;
      const double cy = y0 + ((double)(64 * J) + 31.5) * py;
      g.centre64.push_back(double4{cx, cy, g.z, __builtin_bit_cast(double, rho64)});
      g.cx0_64 = std::min(g.cx0_64, cx);
      g.cx1_64 = std::max(g.cx1_64, cx);
      g.cy0_64 = std::min(g.cy0_64, cy);
      g.cy1_64 = std::max(g.cy1_64, cy);
    }
}

bool build_grid(const double* u, size_t N, HostGrid& g) {
  g = HostGrid();
  if (N < 64) return false;
  size_t ny = 1;
  while (ny < N && u[3 * ny] == u[0]) ++ny;
  if (ny == N || N % ny) return false;
  const size_t nx = N / ny;
  if (nx % 8 || ny % 8) return false;
  const double z = u[2];
  for (size_t n = 0; n < N; ++n) {
    const size_t i = n / ny, j = n % ny;
    if (u[3 * n] != u[3 * i * ny] || u[3 * n + 1] != u[3 * j + 1] || u[3 * n + 2] != z)
      return false;
  }
  g.nx = nx;
  g.ny = ny;
  g.z = z;
  const size_t TI = nx / 8, TJ = ny / 8;
  g.centre.reserve(TI * TJ);
  g.goff.reserve(4 * TI * TJ);
  g.perm.reserve(64 * TI * TJ);
  for (size_t I = 0; I < TI; ++I)
    for (size_t J = 0; J < TJ; ++J) {
      double xs[8], ys[8];
      for (int a = 0; a < 8; ++a) xs[a] = u[3 * (8 * I + a) * ny];
      for (int b = 0; b < 8; ++b) ys[b] = u[3 * (8 * J + b) + 1];
      const double cx = 0.5 * (*std::min_element(xs, xs + 8) + *std::max_element(xs, xs + 8));
      const double cy = 0.5 * (*std::min_element(ys, ys + 8) + *std::max_element(ys, ys + 8));
      float gx[8], gy[8];
      double rx = 0.0, ry = 0.0;
      for (int k = 0; k < 8; ++k) {
        gx[k] = (float)(xs[k] - cx);
        gy[k] = (float)(ys[k] - cy);
        rx = std::max(rx, std::fabs(xs[k] - cx));
        ry = std::max(ry, std::fabs(ys[k] - cy));
      }
      const float2 rho{std::nextafter((float)rx, INFINITY), std::nextafter((float)ry, INFINITY)};
      g.centre.push_back(double4{cx, cy, z, __builtin_bit_cast(double, rho)});
      g.cx0 = std::min(g.cx0, cx);
      g.cx1 = std::max(g.cx1, cx);
      g.cy0 = std::min(g.cy0, cy);
      g.cy1 = std::max(g.cy1, cy);
      g.rho_x = std::max(g.rho_x, rho.x);
      g.rho_y = std::max(g.rho_y, rho.y);
      const float k = 1024.f;  // 2^10: exact
      for (int q = 0; q < 4; ++q)
        g.goff.push_back(float4{k * gx[2 * q], k * gx[2 * q + 1], k * gy[2 * q],
                                k * gy[2 * q + 1]});
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b)
          g.perm.push_back((int32_t)((8 * I + a) * ny + 8 * J + b));
    }
  // uniform lattices also get 16-row tiles, padded where a side is not a
  // multiple of 16 (1000 x 1000: 1008 x 1024 slots, 3 % of them padding, at
  // twice the 8 x 8 kernel's rate)
  build_grid16(u, g);
  return true;
}

// Number of leading complete rows of an x-slow product grid with rows of ny:
// row i (sources i*ny .. i*ny + ny - 1) is complete when all of its sources
// have the x of its first source, the y of row 0's source in the same column
// and the z of source 0.
size_t complete_rows(const double* u, size_t N, size_t ny) {
  const double z = u[2];
  size_t i = 0;
  for (; (i + 1) * ny <= N; ++i) {
    const double x = u[3 * i * ny];
    for (size_t j = 0; j < ny; ++j) {
      const double* p = &u[3 * (i * ny + j)];
      if (p[0] != x || p[1] != u[3 * j + 1] || p[2] != z) return i;
    }
  }
  return i;
}

// A uniform lattice that build_grid does not take: sides of at least 16, not
// both multiples of 8, or followed by fewer than max(nx, ny) other sources at
// any positions (the reference's plane generator fills a point count that is
// not nx * ny with copies of source 0, cuda/init.h:153-157).  Only the 16-row
// tiles are built, over the lattice padded to multiples of 16; every target
// sums the remaining sources (tail) exactly.
bool build_grid16_padded(const double* u, size_t N, HostGrid& g) {
  g = HostGrid();
  if (N < 256) return false;
  size_t ny = 1;
  while (ny < N && u[3 * ny] == u[0]) ++ny;
  if (ny < 16 || ny == N) return false;
  // nx = the number of leading complete rows: row i holds sources
  // i*ny .. i*ny + ny - 1 with one x and the y of row 0's columns.  With an
  // aspect ratio above 1 the generator's copies of point 0 can number ny or
  // more; they repeat row 0's y and so never complete a row.
  const size_t nx = complete_rows(u, N, ny), body = nx * ny;
  if (nx < 16 || N - body >= std::max(nx, ny)) return false;
  g.nx = nx;
  g.ny = ny;
  g.z = u[2];
  build_grid16(u, g);
  if (g.centre16.empty()) return false;
  for (size_t n = body; n < N; ++n) g.tail.push_back((int32_t)n);
  return true;
}

// Builds and uploads the spatial tiling for k_superpose_f32s (u_host: the
// plan's positions on the host, or NULL to read them back from the device).
// Waits for the copies before returning.
holo_status plan_build_tiles(holo_ctx* ctx, const holo_plan* p, const double* u_host) {
  std::vector<double> host;
  if (!u_host && p->N) {
    host.resize(3 * p->N);
    HOLO_HIP(hipMemcpyAsync(host.data(), p->u.ptr, 3 * p->N * sizeof(double),
                            hipMemcpyDeviceToHost, ctx->stream));
    HOLO_HIP(hipStreamSynchronize(ctx->stream));
    u_host = host.data();
  }
  HostTiles t;
  build_tiles(u_host, p->N, p->tile, t);
  const size_t n_tiles = t.centre.size();
  if (n_tiles * (size_t)p->tile > kMaxCount) return fail(HOLO_EINVAL, "too many tile slots");
  const size_t slots = n_tiles * p->tile;
  HOLO_TRY(ensure(p->centre, std::max<size_t>(n_tiles, 1) * sizeof(double4), ctx->device));
  HOLO_TRY(ensure(p->geo, std::max<size_t>(slots, 1) * sizeof(float4), ctx->device));
  HOLO_TRY(ensure(p->perm, std::max<size_t>(slots, 1) * sizeof(int32_t), ctx->device));
  HOLO_TRY(ensure(p->wc, std::max<size_t>(slots, 1) * sizeof(float2), ctx->device));
  if (n_tiles) {
    HOLO_HIP(hipMemcpyAsync(p->centre.ptr, t.centre.data(), n_tiles * sizeof(double4),
                            hipMemcpyHostToDevice, ctx->stream));
    HOLO_HIP(hipMemcpyAsync(p->geo.ptr, t.geo.data(), slots * sizeof(float4),
                            hipMemcpyHostToDevice, ctx->stream));
    HOLO_HIP(hipMemcpyAsync(p->perm.ptr, t.perm.data(), slots * sizeof(int32_t),
                            hipMemcpyHostToDevice, ctx->stream));
  }
  HOLO_HIP(hipStreamSynchronize(ctx->stream));  // host vectors are freed on return
  holo_plan* q = const_cast<holo_plan*>(p);
  q->n_tiles = n_tiles;
  q->max_offset = t.max_offset;
  q->planar = t.planar;
  p->tiles_built = true;
  return HOLO_OK;
}

// The 16-row tiles of grid (and their tail), if any, onto the device.
holo_status upload_grid16(holo_ctx* ctx, holo_plan* p, const HostGrid& grid) {
  const size_t n16 = grid.centre16.size(), nt = grid.tail.size();
  if (!n16) return HOLO_OK;
  if (256 * n16 + nt > kMaxCount) return fail(HOLO_EINVAL, "too many tile slots");
  p->g16_tiles = n16;
  p->g16_tail = nt;
  p->g16_px = grid.px;
  p->g16_py = grid.py;
  p->g16_cx0 = grid.cx0_16;
  p->g16_cx1 = grid.cx1_16;
  p->g16_cy0 = grid.cy0_16;
  p->g16_cy1 = grid.cy1_16;
  p->g16_rho_x = grid.rho16_x;
  p->g16_rho_y = grid.rho16_y;
  p->grid_z = grid.z;
  HOLO_TRY(ensure(p->g16_centre, n16 * sizeof(double4), ctx->device));
  HOLO_TRY(ensure(p->g16_perm, (256 * n16 + nt) * sizeof(int32_t), ctx->device));
  HOLO_TRY(ensure(p->g16_wc, (256 * n16 + nt) * sizeof(float2), ctx->device));
  HOLO_TRY(ensure(p->g16_wq, 6 * 64 * n16 * sizeof(uint4), ctx->device));
  HOLO_TRY(ensure(p->g16_wqn, 4 * 64 * n16 * sizeof(uint4), ctx->device));
  HOLO_TRY(ensure(p->g16_ts, n16 * sizeof(float), ctx->device));
  HOLO_HIP(hipMemcpyAsync(p->g16_centre.ptr, grid.centre16.data(), n16 * sizeof(double4),
                          hipMemcpyHostToDevice, ctx->stream));
  p->g16_tj = grid.tj16;
  p->g16_nx = grid.nx;
  p->g16_ny = grid.ny;
  HOLO_TRY(ensure(p->g16_axes, grid.axes16.size() * sizeof(double), ctx->device));
  HOLO_HIP(hipMemcpyAsync(p->g16_axes.ptr, grid.axes16.data(),
                          grid.axes16.size() * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
  HOLO_HIP(hipMemcpyAsync(p->g16_perm.ptr, grid.perm16.data(), 256 * n16 * sizeof(int32_t),
                          hipMemcpyHostToDevice, ctx->stream));
  if (nt)
    HOLO_HIP(hipMemcpyAsync(static_cast<int32_t*>(p->g16_perm.ptr) + 256 * n16,
                            grid.tail.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice,
                            ctx->stream));
  const size_t n32 = grid.centre32.size();
  if (n32) {
    p->g32_tiles = n32;
    p->g32_cx0 = grid.cx0_32;
    p->g32_cx1 = grid.cx1_32;
    p->g32_cy0 = grid.cy0_32;
    p->g32_cy1 = grid.cy1_32;
    p->g32_rho_y = grid.rho32_y;
    HOLO_TRY(ensure(p->g32_centre, n32 * sizeof(double4), ctx->device));
    HOLO_HIP(hipMemcpyAsync(p->g32_centre.ptr, grid.centre32.data(), n32 * sizeof(double4),
                            hipMemcpyHostToDevice, ctx->stream));
  }
  const size_t n64 = grid.centre64.size();
  if (n64) {
    p->g64_tiles = n64;
    p->g64_cx0 = grid.cx0_64;
    p->g64_cx1 = grid.cx1_64;
    p->g64_cy0 = grid.cy0_64;
    p->g64_cy1 = grid.cy1_64;
    p->g64_rho_y = grid.rho64_y;
    HOLO_TRY(ensure(p->g64_centre, n64 * sizeof(double4), ctx->device));
    HOLO_HIP(hipMemcpyAsync(p->g64_centre.ptr, grid.centre64.data(), n64 * sizeof(double4),
                            hipMemcpyHostToDevice, ctx->stream));
  }
  return HOLO_OK;
}

holo_status plan_upload_tiles(holo_ctx* ctx, holo_plan* p, const double* u_host) {
  HostGrid grid;
  if (p->tile == 64 && build_grid(u_host, p->N, grid)) {
    const size_t nt = grid.centre.size();
    if (nt * 64 > kMaxCount) return fail(HOLO_EINVAL, "too many tile slots");
    p->grid_nx = grid.nx;
    p->grid_ny = grid.ny;
    p->grid_cx0 = grid.cx0;
    p->grid_cx1 = grid.cx1;
    p->grid_cy0 = grid.cy0;
    p->grid_cy1 = grid.cy1;
    p->grid_z = grid.z;
    p->grid_rho_x = grid.rho_x;
    p->grid_rho_y = grid.rho_y;
    p->grid_tiles = nt;
    HOLO_TRY(ensure(p->g_centre, nt * sizeof(double4), ctx->device));
    HOLO_TRY(ensure(p->g_goff, 4 * nt * sizeof(float4), ctx->device));
    HOLO_TRY(ensure(p->g_perm, 64 * nt * sizeof(int32_t), ctx->device));
    HOLO_TRY(ensure(p->g_wc, 64 * nt * sizeof(float2), ctx->device));
    HOLO_TRY(ensure(p->g_wh, 64 * nt * sizeof(uint4), ctx->device));
    HOLO_TRY(ensure(p->g_wl, 64 * nt * sizeof(uint4), ctx->device));
    HOLO_TRY(ensure(p->g_ts, nt * sizeof(float), ctx->device));
    HOLO_HIP(hipMemcpyAsync(p->g_centre.ptr, grid.centre.data(), nt * sizeof(double4),
                            hipMemcpyHostToDevice, ctx->stream));
    HOLO_HIP(hipMemcpyAsync(p->g_goff.ptr, grid.goff.data(), 4 * nt * sizeof(float4),
                            hipMemcpyHostToDevice, ctx->stream));
    HOLO_HIP(hipMemcpyAsync(p->g_perm.ptr, grid.perm.data(), 64 * nt * sizeof(int32_t),
                            hipMemcpyHostToDevice, ctx->stream));
    HOLO_TRY(upload_grid16(ctx, p, grid));
  } else if (p->tile == 64 && build_grid16_padded(u_host, p->N, grid)) {
    p->grid_nx = grid.nx;
    p->grid_ny = grid.ny;
    HOLO_TRY(upload_grid16(ctx, p, grid));
  }
  // host vectors are freed on return: wait for the copies
  HOLO_HIP(hipStreamSynchronize(ctx->stream));
  if (!p->grid_tiles) return plan_build_tiles(ctx, p, u_host);
  // a lattice: the same 8 x 8 tiles the spatial tiler would cut (built on demand)
  p->n_tiles = p->grid_tiles;
  p->max_offset = std::hypot((double)p->grid_rho_x, (double)p->grid_rho_y);
  p->planar = true;
  return HOLO_OK;
}

// Every source has the z of source 0 (== on the values: +0 and -0 give the
// same dz bits up to sign, and dz enters only as dz^2 and |dz|).
bool uniform_z(const double* u, size_t N) {
  for (size_t n = 1; n < N; ++n)
    if (!(u[3 * n + 2] == u[2])) return false;
  return N > 0;
}

// Length of the rows the sources form: runs of `row` consecutive sources with
// one x, the last run possibly shorter (the x-slow lattices of
// cuda/init.h:89-158 give row = ny, and the copies of source 0 that fill a
// plane whose point count is not nx * ny form the short last run).  The run
// length of source 0's x is the candidate; 0 when that is under 4 sources or
// when any source differs in x from the first of its row.
size_t row_length(const double* u, size_t N) {
  size_t L = 1;
  while (L < N && u[3 * L] == u[0]) ++L;
  if (L < 4) return 0;
  for (size_t n = L; n < N; ++n)
    if (!(u[3 * n] == u[3 * (n - n % L)])) return 0;
  return L;
}

holo_status plan_build(holo_ctx* ctx, holo_plan* p, const double* u, size_t N,
                       unsigned tile, bool dev) {
  p->device = ctx->device;
  p->N = N;
  p->tile = tile ? tile : 64;
  if (p->tile != 8 && p->tile != 16 && p->tile != 32 && p->tile != 64)
    return fail(HOLO_EINVAL, "tile must be 8, 16, 32 or 64");
  if (N && !u) return fail(HOLO_EINVAL, "null source positions");
  std::vector<double> host;
  const double* uh = u;
  HOLO_TRY(ensure(p->u, std::max<size_t>(3 * N, 1) * sizeof(double), ctx->device));
  if (N) {
    if (dev) {
      host.resize(3 * N);
      HOLO_HIP(hipMemcpyAsync(host.data(), u, 3 * N * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
      HOLO_HIP(hipStreamSynchronize(ctx->stream));
      uh = host.data();
      HOLO_HIP(hipMemcpyAsync(p->u.ptr, u, 3 * N * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
    } else {
      HOLO_HIP(hipMemcpyAsync(p->u.ptr, u, 3 * N * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
    }
  }
  p->uniform_z = uniform_z(uh, N);
  p->row = row_length(uh, N);
  HOLO_TRY(plan_upload_tiles(ctx, p, uh));
  // the uploads are ordered before a first call on another stream
  HOLO_HIP(hipEventCreateWithFlags(&p->ran, hipEventDisableTiming));
  HOLO_HIP(hipEventRecord(p->ran, ctx->stream));
  p->ran_on = ctx->stream;
  return HOLO_OK;
}

// y (device) = sum over the plan's sources; weights already device-resident.
holo_status plan_run_body(holo_ctx* ctx, const holo_plan* p, int precision, int sign,
                          bool obliquity, double lambda, holo::WeightForm wf,
                          const double* w0, const double* w1, const double* v, size_t M,
                          double* y) {
  if (precision == HOLO_FP64) {
    const size_t pe = split_elems_f64(M, p->N);
    if (pe) HOLO_TRY(ensure(ctx->partial, pe * sizeof(double2), ctx->device));
    const double K = (2 * M_PI) / lambda;  // TWO_PI_OVER_LAMBDA, cuda/macros.h:29-31
    HOLO_HIP(holo::launch_superpose_f64(sign, wf, obliquity, p->uniform_z, p->row, w0, w1,
                                        static_cast<const double*>(p->u.ptr),
                                        p->N, v, M, K, reinterpret_cast<double2*>(y),
                                        static_cast<double2*>(ctx->partial.ptr), pe,
                                        ctx->stream));
  } else if (p->g16_tiles && holo::tuning().grid && holo::tuning().grid16) {
    const size_t nt = p->g16_tiles;
    const bool pairs = p->g32_tiles && holo::tuning().grid32;
    const bool quads = pairs && p->g64_tiles && holo::tuning().grid64;
    const int nkc = quads ? 4 : (pairs ? 2 : 1);
    HOLO_HIP(holo::launch_tile_weights(wf, w0, w1, static_cast<const int32_t*>(p->g16_perm.ptr),
                                       256 * nt + p->g16_tail,
                                       static_cast<float2*>(p->g16_wc.ptr), ctx->stream));
    HOLO_HIP(holo::launch_grid16_weights(static_cast<const float2*>(p->g16_wc.ptr), (uint32_t)nt,
                                         nkc, static_cast<uint4*>(p->g16_wq.ptr),
                                         static_cast<uint4*>(p->g16_wqn.ptr),
                                         static_cast<float*>(p->g16_ts.ptr), ctx->stream));
    holo::GridTiles g16;  // the 16 x 16 tiles
    g16.perm = static_cast<const int32_t*>(p->g16_perm.ptr);
    g16.u = static_cast<const double*>(p->u.ptr);
    g16.z = p->grid_z;
    g16.rho_x = p->g16_rho_x;
    g16.px = p->g16_px;
    g16.py = p->g16_py;
    g16.centre = static_cast<const double4*>(p->g16_centre.ptr);
    g16.n_tiles = (uint32_t)nt;
    g16.cx0 = p->g16_cx0;
    g16.cx1 = p->g16_cx1;
    g16.cy0 = p->g16_cy0;
    g16.cy1 = p->g16_cy1;
    g16.rho_y = p->g16_rho_y;
    g16.tail_perm = g16.perm + 256 * nt;
    g16.tail_wc = static_cast<const float2*>(p->g16_wc.ptr) + 256 * nt;
    g16.n_tail = (uint32_t)p->g16_tail;
    if (holo::tuning().grid_rows) {
      g16.xs = static_cast<const double*>(p->g16_axes.ptr);
      g16.ys = g16.xs + 16 * (nt / p->g16_tj);
    }
    g16.tj = (uint32_t)p->g16_tj;
    g16.nx = (uint32_t)p->g16_nx;
    g16.ny = (uint32_t)p->g16_ny;
    holo::GridTiles g32 = g16;  // the pairs of them
    if (pairs) {
      g32.centre = static_cast<const double4*>(p->g32_centre.ptr);
      g32.n_tiles = (uint32_t)p->g32_tiles;
      g32.cx0 = p->g32_cx0;
      g32.cx1 = p->g32_cx1;
      g32.cy0 = p->g32_cy0;
      g32.cy1 = p->g32_cy1;
      g32.rho_y = p->g32_rho_y;
    }
    holo::GridTiles gt = pairs ? g32 : g16;  // the tiles of the launch
    if (quads) {
      gt.centre = static_cast<const double4*>(p->g64_centre.ptr);
      gt.n_tiles = (uint32_t)p->g64_tiles;
      gt.cx0 = p->g64_cx0;
      gt.cx1 = p->g64_cx1;
      gt.cy0 = p->g64_cy0;
      gt.cy1 = p->g64_cy1;
      gt.rho_y = p->g64_rho_y;
    }
    const int S = holo::grid_splits(M, gt.n_tiles);
    const size_t pe = S > 1 ? (size_t)S * M : 0;
    if (pe) HOLO_TRY(ensure(ctx->partial, pe * sizeof(double2), ctx->device));
    HOLO_HIP(holo::launch_superpose_grid16(sign, obliquity, nkc, gt, g32, g16,
                                           static_cast<const float2*>(p->g16_wc.ptr),
                                           static_cast<const uint4*>(p->g16_wq.ptr),
                                           static_cast<const uint4*>(p->g16_wqn.ptr),
                                           static_cast<const float*>(p->g16_ts.ptr), v, M,
                                           lambda, reinterpret_cast<double2*>(y),
                                           static_cast<double2*>(ctx->partial.ptr), pe,
                                           ctx->stream));
  } else if (p->grid_tiles && holo::tuning().grid) {
    const size_t nt = p->grid_tiles;
    HOLO_HIP(holo::launch_tile_weights(wf, w0, w1, static_cast<const int32_t*>(p->g_perm.ptr),
                                       64 * nt, static_cast<float2*>(p->g_wc.ptr), ctx->stream));
    HOLO_HIP(holo::launch_grid_weights(static_cast<const float2*>(p->g_wc.ptr),
                                       static_cast<const float4*>(p->g_goff.ptr), (uint32_t)nt,
                                       static_cast<uint4*>(p->g_wh.ptr),
                                       static_cast<uint4*>(p->g_wl.ptr),
                                       static_cast<float*>(p->g_ts.ptr), ctx->stream));
    holo::GridTiles gt;
    gt.centre = static_cast<const double4*>(p->g_centre.ptr);
    gt.goff = static_cast<const float4*>(p->g_goff.ptr);
    gt.perm = static_cast<const int32_t*>(p->g_perm.ptr);
    gt.u = static_cast<const double*>(p->u.ptr);
    gt.n_tiles = (uint32_t)nt;
    gt.cx0 = p->grid_cx0;
    gt.cx1 = p->grid_cx1;
    gt.cy0 = p->grid_cy0;
    gt.cy1 = p->grid_cy1;
    gt.z = p->grid_z;
    gt.rho_x = p->grid_rho_x;
    gt.rho_y = p->grid_rho_y;
    const int S = holo::grid_splits(M, nt);
    const size_t pe = S > 1 ? (size_t)S * M : 0;
    if (pe) HOLO_TRY(ensure(ctx->partial, pe * sizeof(double2), ctx->device));
    HOLO_HIP(holo::launch_superpose_grid(sign, obliquity, gt,
                                         static_cast<const float2*>(p->g_wc.ptr),
                                         static_cast<const uint4*>(p->g_wh.ptr),
                                         static_cast<const uint4*>(p->g_wl.ptr),
                                         static_cast<const float*>(p->g_ts.ptr), v, M, lambda,
                                         reinterpret_cast<double2*>(y),
                                         static_cast<double2*>(ctx->partial.ptr), pe,
                                         ctx->stream));
  } else {
    {
      std::lock_guard<std::mutex> lock(p->tiles_mu);
      if (!p->tiles_built) HOLO_TRY(plan_build_tiles(ctx, p, nullptr));
    }
    const size_t slots = p->n_tiles * p->tile;
    HOLO_HIP(holo::launch_tile_weights(wf, w0, w1, static_cast<const int32_t*>(p->perm.ptr),
                                       slots, static_cast<float2*>(p->wc.ptr), ctx->stream,
                                       true));
    holo::TileSet ts;
    ts.centre = static_cast<const double4*>(p->centre.ptr);
    ts.geo = static_cast<const float4*>(p->geo.ptr);
    ts.perm = static_cast<const int32_t*>(p->perm.ptr);
    ts.u = static_cast<const double*>(p->u.ptr);
    ts.n_tiles = (uint32_t)p->n_tiles;
    ts.tile = p->tile;
    ts.planar = p->planar;
    const size_t pe = split_elems_f32(M, p->n_tiles);
    if (pe) HOLO_TRY(ensure(ctx->partial, pe * sizeof(double2), ctx->device));
    HOLO_HIP(holo::launch_superpose_f32s(sign, obliquity, ts,
                                         static_cast<const float2*>(p->wc.ptr), v, M,
                                         lambda, reinterpret_cast<double2*>(y),
                                         static_cast<double2*>(ctx->partial.ptr), pe,
                                         ctx->stream));
  }
  return HOLO_OK;
}

// A plan may serve contexts on different streams.  A call on a stream other
// than the one of the plan's previous call waits on the device (not the host)
// until that call has finished with the plan's buffers.
holo_status plan_run(holo_ctx* ctx, const holo_plan* p, int precision, int sign,
                     bool obliquity, double lambda, holo::WeightForm wf, const double* w0,
                     const double* w1, const double* v, size_t M, double* y) {
  if (M == 0) return HOLO_OK;
  if (p->ran && p->ran_on != ctx->stream) HOLO_HIP(hipStreamWaitEvent(ctx->stream, p->ran, 0));
  const holo_status s = plan_run_body(ctx, p, precision, sign, obliquity, lambda, wf, w0, w1,
                                      v, M, y);
  // recorded after a failed call too: some of its kernels may be queued
  if (!p->ran) HOLO_HIP(hipEventCreateWithFlags(&p->ran, hipEventDisableTiming));
  HOLO_HIP(hipEventRecord(p->ran, ctx->stream));
  p->ran_on = ctx->stream;
  return s;
}

// The FP64 formulation over the plan's sources (weights w0, w1 in form wf) at
// the error report's subset: e.ref[i] for the count targets e.vsub.
holo_status plan_fp64_at(holo_ctx* ctx, const holo_plan* p, int sign, bool obliquity,
                         double lambda, holo::WeightForm wf, const double* w0, const double* w1,
                         const ErrScratch& e, size_t count) {
  const double K = (2 * M_PI) / lambda;  // TWO_PI_OVER_LAMBDA, cuda/macros.h:29-31
  HOLO_HIP(holo::launch_superpose_f64(sign, wf, obliquity, p->uniform_z, p->row, w0, w1,
                                      static_cast<const double*>(p->u.ptr), p->N, e.vsub, count,
                                      K, e.ref, e.partial, e.partial_cap, ctx->stream, true));
  return HOLO_OK;
}

}  // namespace

// ===========================================================================

extern "C" {

int holo_abi_version(void) { return HOLO_ABI_VERSION; }

const char* holo_status_string(holo_status s) {
  switch (s) {
    case HOLO_OK: return "ok";
    case HOLO_EINVAL: return "invalid argument";
    case HOLO_ENOMEM: return "out of memory";
    case HOLO_EDEVICE: return "device error";
    case HOLO_EIO: return "i/o error";
  }
  return "unknown status";
}

const char* holo_last_error(void) { return g_error.c_str(); }

holo_status holo_create(holo_ctx** out, int device, void* stream) {
  if (!out) return fail(HOLO_EINVAL, "null context pointer");
  *out = nullptr;
  int count = 0;
  HOLO_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(HOLO_EDEVICE, "no HIP device " + std::to_string(device));
  hipDeviceProp_t prop;
  HOLO_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(HOLO_EDEVICE, std::string("libholo is built for gfx950, device is ") +
                                  prop.gcnArchName);
  HOLO_HIP(hipSetDevice(device));
  void* scratch_ptr = nullptr;  // the no-allocation entry point's scratch, looked up once
  HOLO_TRY(noalloc_scratch(device, &scratch_ptr));
  holo_ctx* c = new holo_ctx();
  c->device = device;
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fail(HOLO_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  *out = c;
  return HOLO_OK;
}

void holo_destroy(holo_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->switched) (void)hipEventDestroy(ctx->switched);
  delete ctx;
}

holo_status holo_set_stream(holo_ctx* ctx, void* stream) {
  HOLO_TRY(use_device(ctx));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (ctx->stream == s) return HOLO_OK;
  if (ctx->own_stream) {
    HOLO_HIP(hipStreamSynchronize(ctx->stream));
    HOLO_HIP(hipStreamDestroy(ctx->stream));
    ctx->own_stream = false;
  } else {
    // work queued on the previous stream may still use the context's scratch
    // buffers: work queued on `s` from now on waits for it on the device
    if (!ctx->switched)
      HOLO_HIP(hipEventCreateWithFlags(&ctx->switched, hipEventDisableTiming));
    HOLO_HIP(hipEventRecord(ctx->switched, ctx->stream));
    HOLO_HIP(hipStreamWaitEvent(s, ctx->switched, 0));
  }
  ctx->stream = s;
  return HOLO_OK;
}

holo_status holo_synchronize(holo_ctx* ctx) {
  HOLO_TRY(use_device(ctx));
  HOLO_HIP(hipStreamSynchronize(ctx->stream));
  return HOLO_OK;
}

holo_status holo_set_option(const char* name, long value) {
  if (!name) return fail(HOLO_EINVAL, "null option name");
  holo::Tuning& t = holo::tuning();
  const std::string n(name);
  if (n == "r_f64" || n == "r_f32") {
    if (value != 1 && value != 2 && value != 4)
      return fail(HOLO_EINVAL, n + " must be 1, 2 or 4");
    (n == "r_f64" ? t.r_f64 : t.r_f32) = (int)value;
  } else if (n == "splits") {
    if (value < 0 || value > 64) return fail(HOLO_EINVAL, "splits must be in [0, 64]");
    t.splits = (int)value;
  } else if (n == "target_waves") {
    if (value < 1) return fail(HOLO_EINVAL, "target_waves must be positive");
    t.target_waves = (int)value;
  } else if (n == "f64_libm") {
    t.f64_libm = value != 0;
  } else if (n == "f64_planar") {
    t.f64_planar = value != 0;
  } else if (n == "f64_rows") {
    t.f64_rows = value != 0;
  } else if (n == "grid") {
    t.grid = value != 0;
  } else if (n == "grid16") {
    t.grid16 = value != 0;
  } else if (n == "grid32") {
    t.grid32 = value != 0;
  } else if (n == "grid64") {
    t.grid64 = value != 0;
  } else if (n == "grid_series") {
    t.grid_series = value != 0;
  } else if (n == "grid_near") {
    t.grid_near = value != 0;
  } else if (n == "grid_rows") {
    t.grid_rows = value != 0;
  } else if (n == "mc_lattice") {
    t.mc_lattice = value != 0;
  } else if (n == "mc_passes") {
    if (value < 0 || value > 64) return fail(HOLO_EINVAL, "mc_passes must be in [0, 64]");
    t.mc_passes = (int)value;
  } else {
    return fail(HOLO_EINVAL, "unknown option " + n);
  }
  return HOLO_OK;
}

long holo_get_option(const char* name) {
  if (!name) return -1;
  const holo::Tuning& t = holo::tuning();
  const std::string n(name);
  if (n == "r_f64") return t.r_f64;
  if (n == "r_f32") return t.r_f32;
  if (n == "splits") return t.splits;
  if (n == "target_waves") return t.target_waves;
  if (n == "f64_libm") return t.f64_libm ? 1 : 0;
  if (n == "f64_planar") return t.f64_planar ? 1 : 0;
  if (n == "f64_rows") return t.f64_rows ? 1 : 0;
  if (n == "grid") return t.grid ? 1 : 0;
  if (n == "grid16") return t.grid16 ? 1 : 0;
  if (n == "grid32") return t.grid32 ? 1 : 0;
  if (n == "grid64") return t.grid64 ? 1 : 0;
  if (n == "grid_series") return t.grid_series ? 1 : 0;
  if (n == "grid_near") return t.grid_near ? 1 : 0;
  if (n == "grid_rows") return t.grid_rows ? 1 : 0;
  if (n == "mc_lattice") return t.mc_lattice ? 1 : 0;
  if (n == "mc_passes") return t.mc_passes;
  return -1;
}

holo_status holo_transform_full(holo_ctx* ctx, holo_precision precision, int sign,
                                double lambda, const double* amp, const double* phase,
                                const double* u, size_t N, const double* v, size_t M,
                                double* y, double* err_rel_l2, unsigned flags) {
  HOLO_TRY(check_common(ctx, precision, sign, lambda, N, M));
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS | HOLO_OBLIQUITY));
  HOLO_TRY(use_device(ctx));
  const bool dev = flags & HOLO_DEVICE_POINTERS;
  const bool obl = flags & HOLO_OBLIQUITY;
  if (M == 0) return HOLO_OK;
  double* dy;
  HOLO_TRY(out_buf(ctx, ctx->y, y, 2 * M, dev, &dy));
  const double *da, *dp, *dv;
  HOLO_TRY(stage_in(ctx, ctx->amp, amp, N, dev, &da));
  HOLO_TRY(stage_in(ctx, ctx->phase, phase, N, dev, &dp));
  HOLO_TRY(stage_in(ctx, ctx->v, v, 3 * M, dev, &dv));
  if (N == 0) {
    HOLO_HIP(holo::launch_zero(dy, 2 * M, ctx->stream));
    HOLO_TRY(report_zero(ctx, err_rel_l2, dev));
    return finish_out(ctx, y, dy, 2 * M, dev);
  }
  if (precision == HOLO_FP64) {
    HOLO_TRY(report_zero(ctx, err_rel_l2, dev));
    const double* du;
    HOLO_TRY(stage_in(ctx, ctx->u, u, 3 * N, dev, &du));
    const size_t pe = split_elems_f64(M, N);
    if (pe) HOLO_TRY(ensure(ctx->partial, pe * sizeof(double2), ctx->device));
    const double K = (2 * M_PI) / lambda;  // TWO_PI_OVER_LAMBDA, cuda/macros.h:29-31
    if (dev && !holo::tuning().f64_libm) {
      // device positions: their layout is found on the device
      HOLO_TRY(ensure(ctx->layout, 4 * sizeof(uint32_t), ctx->device));
      HOLO_HIP(holo::launch_superpose_f64_auto(
          sign, holo::kPolar, obl, da, dp, du, N, dv, M, K, reinterpret_cast<double2*>(dy),
          static_cast<double2*>(ctx->partial.ptr), pe, static_cast<uint32_t*>(ctx->layout.ptr),
          ctx->stream));
    } else {
      const bool uz = !dev && uniform_z(u, N);
      const size_t row = uz ? row_length(u, N) : 0;
      HOLO_HIP(holo::launch_superpose_f64(sign, holo::kPolar, obl, uz, row, da, dp, du, N, dv,
                                          M, K, reinterpret_cast<double2*>(dy),
                                          static_cast<double2*>(ctx->partial.ptr), pe,
                                          ctx->stream));
    }
  } else {
    holo_plan plan;
    HOLO_TRY(plan_build(ctx, &plan, u, N, 0, dev));
    HOLO_TRY(plan_run(ctx, &plan, precision, sign, obl, lambda, holo::kPolar, da, dp, dv, M,
                      dy));
    HOLO_TRY(report_error(ctx, false, dv, M, dy, err_rel_l2, dev,
                          [&](const ErrScratch& e, size_t count) {
                            return plan_fp64_at(ctx, &plan, sign, obl, lambda, holo::kPolar, da,
                                                dp, e, count);
                          }));
    HOLO_HIP(hipStreamSynchronize(ctx->stream));  // plan buffers die here
  }
  return finish_out(ctx, y, dy, 2 * M, dev);
}

holo_status holo_transform_full_noalloc(int sign, double lambda, const double* amp,
                                        const double* phase, const double* u, size_t N,
                                        const double* v, size_t M, double* y, void* stream,
                                        unsigned flags) {
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS | HOLO_OBLIQUITY));
  if (sign != 1 && sign != -1) return fail(HOLO_EINVAL, "sign must be +1 or -1");
  if (!(lambda > 0.0) || !std::isfinite(lambda))
    return fail(HOLO_EINVAL, "lambda must be positive and finite");
  if (N > kMaxCount || M > kMaxCount) return fail(HOLO_EINVAL, "N and M must be below 2^31");
  if (M == 0) return HOLO_OK;
  if (!v || !y) return fail(HOLO_EINVAL, "null target or output array");
  if (N && (!amp || !phase || !u)) return fail(HOLO_EINVAL, "null source array");
  int device = 0;
  HOLO_HIP(hipGetDevice(&device));
  void* base = nullptr;
  HOLO_TRY(noalloc_scratch(device, &base));
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (N == 0) {
    HOLO_HIP(holo::launch_zero(y, 2 * M, st));
    return HOLO_OK;
  }
  uint32_t* layout = static_cast<uint32_t*>(base);
  double2* partial =
      reinterpret_cast<double2*>(static_cast<unsigned char*>(base) + kNoallocLayoutBytes);
  const size_t cap = (kNoallocScratchBytes - kNoallocLayoutBytes) / sizeof(double2);
  const double K = (2 * M_PI) / lambda;  // TWO_PI_OVER_LAMBDA, cuda/macros.h:29-31
  HOLO_HIP(holo::launch_superpose_f64_auto(sign, holo::kPolar, flags & HOLO_OBLIQUITY, amp,
                                           phase, u, N, v, M, K, reinterpret_cast<double2*>(y),
                                           partial, cap, layout, st));
  return HOLO_OK;
}

holo_status holo_transform_mc(holo_ctx* ctx, holo_precision precision, int sign,
                              double lambda, const double* amp, const double* phase,
                              const double* u, size_t N, const double* v, size_t M,
                              uint32_t K, uint64_t seed, double* y, double* err_rel_l2,
                              unsigned flags) {
  HOLO_TRY(check_common(ctx, precision, sign, lambda, N, M));
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS | HOLO_MC_SHARED | HOLO_OBLIQUITY));
  HOLO_TRY(use_device(ctx));
  if (K == 0) return fail(HOLO_EINVAL, "K must be positive");
  if (N == 0) return fail(HOLO_EINVAL, "Monte Carlo needs at least one source");
  const bool dev = flags & HOLO_DEVICE_POINTERS;
  if (M == 0) return HOLO_OK;
  double* dy;
  HOLO_TRY(out_buf(ctx, ctx->y, y, 2 * M, dev, &dy));
  const double *da, *dp, *du, *dv;
  HOLO_TRY(stage_in(ctx, ctx->amp, amp, N, dev, &da));
  HOLO_TRY(stage_in(ctx, ctx->phase, phase, N, dev, &dp));
  HOLO_TRY(stage_in(ctx, ctx->u, u, 3 * N, dev, &du));
  HOLO_TRY(stage_in(ctx, ctx->v, v, 3 * M, dev, &dv));
  const bool fp64 = precision == HOLO_FP64;
  HOLO_TRY(ensure(ctx->pos, N * sizeof(double4), ctx->device));
  HOLO_TRY(ensure(ctx->ph, N * sizeof(double), ctx->device));
  // per-target FP32-split draws on product-grid sources gather only the
  // weight; the grid is recognised on the device, without a host round trip
  const bool lattice = !fp64 && !(flags & HOLO_MC_SHARED) && holo::tuning().mc_lattice;
  holo::McLattice lat;
  if (lattice) {
    HOLO_TRY(ensure(ctx->mc_wc, N * sizeof(float2), ctx->device));
    HOLO_TRY(ensure(ctx->mc_xy, (holo::kLatticeMaxRow + 1 + N) * sizeof(double), ctx->device));
    HOLO_TRY(ensure(ctx->mc_info, 2 * sizeof(uint32_t), ctx->device));
    lat.info = static_cast<const uint32_t*>(ctx->mc_info.ptr);
    lat.wc = static_cast<const float2*>(ctx->mc_wc.ptr);
    lat.xy = static_cast<const double*>(ctx->mc_xy.ptr);
  }
  HOLO_HIP(holo::launch_mc_pack(fp64, da, dp, du, N, static_cast<double4*>(ctx->pos.ptr),
                                static_cast<double*>(ctx->ph.ptr),
                                lattice ? static_cast<float2*>(ctx->mc_wc.ptr) : nullptr,
                                ctx->stream));
  if (lattice)
    HOLO_HIP(holo::launch_mc_lattice_probe(du, N, static_cast<uint32_t*>(ctx->mc_info.ptr),
                                           static_cast<double*>(ctx->mc_xy.ptr), ctx->stream));
  const double Kw = (2 * M_PI) / lambda;
  const bool shared = flags & HOLO_MC_SHARED, obl = flags & HOLO_OBLIQUITY;
  HOLO_HIP(holo::launch_mc(sign, fp64, shared, obl, static_cast<const double4*>(ctx->pos.ptr),
                           static_cast<const double*>(ctx->ph.ptr), N, dv, M, K, seed, Kw,
                           lambda, lat, reinterpret_cast<double2*>(dy), ctx->stream));
  // the FP64 estimator with the same indices at the report's targets
  HOLO_TRY(report_error(ctx, fp64, dv, M, dy, err_rel_l2, dev,
                        [&](const ErrScratch& e, size_t count) {
                          HOLO_TRY(ensure(ctx->pos64, N * sizeof(double4), ctx->device));
                          HOLO_TRY(ensure(ctx->ph64, N * sizeof(double), ctx->device));
                          double4* p64 = static_cast<double4*>(ctx->pos64.ptr);
                          double* ph64 = static_cast<double*>(ctx->ph64.ptr);
                          HOLO_HIP(holo::launch_mc_pack(true, da, dp, du, N, p64, ph64, nullptr,
                                                        ctx->stream));
                          HOLO_HIP(holo::launch_mc_subset(sign, shared, obl, p64, ph64, N,
                                                          e.vsub, count, e.tmap, K, seed, Kw,
                                                          lambda, e.ref, e.partial,
                                                          e.partial_cap, ctx->stream));
                          return HOLO_OK;
                        }));
  return finish_out(ctx, y, dy, 2 * M, dev);
}

holo_status holo_mc_indices(holo_ctx* ctx, uint64_t seed, uint64_t N, size_t m0,
                            size_t count, uint32_t K, int64_t* idx, unsigned flags) {
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS | HOLO_MC_SHARED));
  HOLO_TRY(use_device(ctx));
  if (N == 0 || N > kMaxCount) return fail(HOLO_EINVAL, "N must be in [1, 2^31)");
  if (m0 + count > kMaxCount) return fail(HOLO_EINVAL, "target range beyond 2^31");
  const bool dev = flags & HOLO_DEVICE_POINTERS;
  const size_t total = count * (size_t)K;
  if (!total) return HOLO_OK;
  int64_t* di;
  HOLO_TRY(out_buf(ctx, ctx->w, idx, total, dev, &di));
  HOLO_HIP(holo::launch_mc_indices(seed, N, m0, count, K, flags & HOLO_MC_SHARED, di,
                                   ctx->stream));
  return finish_out(ctx, idx, di, total, dev);
}

holo_status holo_normalize_polar(holo_ctx* ctx, const double* y, size_t M, double* amp,
                                 double* phase, double* max_amp, unsigned flags) {
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS));
  HOLO_TRY(use_device(ctx));
  if (M > kMaxCount) return fail(HOLO_EINVAL, "M must be below 2^31");
  const bool dev = flags & HOLO_DEVICE_POINTERS;
  if (M == 0) {
    if (max_amp && !dev) *max_amp = 0.0;
    return HOLO_OK;
  }
  const double* dy;
  HOLO_TRY(stage_in(ctx, ctx->y, y, 2 * M, dev, &dy));
  double *da, *dp;
  HOLO_TRY(out_buf(ctx, ctx->amp, amp, M, dev, &da));
  HOLO_TRY(out_buf(ctx, ctx->phase, phase, M, dev, &dp));
  HOLO_TRY(ensure(ctx->small, 2048 * sizeof(double), ctx->device));
  double* scratch = static_cast<double*>(ctx->small.ptr);
  double* dmax = (dev && max_amp) ? max_amp : scratch + 1024;
  HOLO_HIP(holo::launch_normalize_polar(reinterpret_cast<const double2*>(dy), M, da, dp, dmax,
                                        scratch, ctx->stream));
  if (!dev) {
    HOLO_HIP(hipMemcpyAsync(amp, da, M * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HOLO_HIP(hipMemcpyAsync(phase, dp, M * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    double mx = 0.0;
    HOLO_HIP(hipMemcpyAsync(&mx, dmax, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HOLO_HIP(hipStreamSynchronize(ctx->stream));
    if (max_amp) *max_amp = mx;
  }
  return HOLO_OK;
}

holo_status holo_plan_create(holo_ctx* ctx, holo_plan** plan, const double* u, size_t N,
                             unsigned tile, unsigned flags) {
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS));
  HOLO_TRY(use_device(ctx));
  if (!plan) return fail(HOLO_EINVAL, "null plan pointer");
  *plan = nullptr;
  if (N > kMaxCount) return fail(HOLO_EINVAL, "N must be below 2^31");
  holo_plan* p = new holo_plan();
  holo_status s = plan_build(ctx, p, u, N, tile, flags & HOLO_DEVICE_POINTERS);
  if (s != HOLO_OK) {
    delete p;
    return s;
  }
  *plan = p;
  return HOLO_OK;
}

void holo_plan_destroy(holo_plan* plan) { delete plan; }

holo_status holo_plan_lattice(const holo_plan* plan, size_t* nx, size_t* ny) {
  if (!plan) return fail(HOLO_EINVAL, "null plan");
  if (nx) *nx = plan->grid_nx;
  if (ny) *ny = plan->grid_ny;
  return HOLO_OK;
}

holo_status holo_plan_info(const holo_plan* plan, size_t* n_sources, size_t* n_tiles,
                           unsigned* tile, double* max_offset) {
  if (!plan) return fail(HOLO_EINVAL, "null plan");
  if (n_sources) *n_sources = plan->N;
  if (n_tiles) *n_tiles = plan->n_tiles;
  if (tile) *tile = plan->tile;
  if (max_offset) *max_offset = plan->max_offset;
  return HOLO_OK;
}

static holo_status plan_transform_any(holo_ctx* ctx, const holo_plan* plan, int precision,
                                      int sign, double lambda, holo::WeightForm wf,
                                      const double* w0, const double* w1, const double* v,
                                      size_t M, double* y, double* err, unsigned flags) {
  if (!plan) return fail(HOLO_EINVAL, "null plan");
  HOLO_TRY(check_common(ctx, precision, sign, lambda, plan->N, M));
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS | HOLO_OBLIQUITY));
  HOLO_TRY(use_device(ctx));
  if (plan->device != ctx->device) return fail(HOLO_EINVAL, "plan belongs to another device");
  const bool dev = flags & HOLO_DEVICE_POINTERS;
  if (M == 0) return HOLO_OK;
  double* dy;
  HOLO_TRY(out_buf(ctx, ctx->y, y, 2 * M, dev, &dy));
  const double* dv;
  HOLO_TRY(stage_in(ctx, ctx->v, v, 3 * M, dev, &dv));
  const size_t N = plan->N;
  if (N == 0) {
    HOLO_HIP(holo::launch_zero(dy, 2 * M, ctx->stream));
    HOLO_TRY(report_zero(ctx, err, dev));
    return finish_out(ctx, y, dy, 2 * M, dev);
  }
  const double *d0, *d1 = nullptr;
  if (wf == holo::kPolar) {
    HOLO_TRY(stage_in(ctx, ctx->amp, w0, N, dev, &d0));
    HOLO_TRY(stage_in(ctx, ctx->phase, w1, N, dev, &d1));
  } else {
    HOLO_TRY(stage_in(ctx, ctx->amp, w0, 2 * N, dev, &d0));
  }
  const bool obl = flags & HOLO_OBLIQUITY;
  HOLO_TRY(plan_run(ctx, plan, precision, sign, obl, lambda, wf, d0, d1, dv, M, dy));
  HOLO_TRY(report_error(ctx, precision == HOLO_FP64, dv, M, dy, err, dev,
                        [&](const ErrScratch& e, size_t count) {
                          return plan_fp64_at(ctx, plan, sign, obl, lambda, wf, d0, d1, e,
                                              count);
                        }));
  return finish_out(ctx, y, dy, 2 * M, dev);
}

holo_status holo_plan_transform(holo_ctx* ctx, const holo_plan* plan, holo_precision precision,
                                int sign, double lambda, const double* amp,
                                const double* phase, const double* v, size_t M, double* y,
                                double* err_rel_l2, unsigned flags) {
  return plan_transform_any(ctx, plan, precision, sign, lambda, holo::kPolar, amp, phase, v, M,
                            y, err_rel_l2, flags);
}

holo_status holo_plan_transform_complex(holo_ctx* ctx, const holo_plan* plan,
                                        holo_precision precision, int sign, double lambda,
                                        const double* w, const double* v, size_t M, double* y,
                                        double* err_rel_l2, unsigned flags) {
  return plan_transform_any(ctx, plan, precision, sign, lambda, holo::kCartesian, w, nullptr,
                            v, M, y, err_rel_l2, flags);
}

holo_status holo_plan_adjoint(holo_ctx* ctx, const holo_plan* plan_targets,
                              holo_precision precision, int sign, double lambda,
                              const double* g, const double* amp, const double* theta,
                              const double* u, size_t N, double* grad_amp, double* grad_theta,
                              double* err_rel_l2, unsigned flags) {
  if (!plan_targets) return fail(HOLO_EINVAL, "null plan");
  HOLO_TRY(check_common(ctx, precision, sign, lambda, N, plan_targets->N));
  HOLO_TRY(check_flags(flags, HOLO_DEVICE_POINTERS | HOLO_OBLIQUITY));
  HOLO_TRY(use_device(ctx));
  if (plan_targets->device != ctx->device)
    return fail(HOLO_EINVAL, "plan belongs to another device");
  const bool dev = flags & HOLO_DEVICE_POINTERS;
  if (N == 0) return report_zero(ctx, err_rel_l2, dev);
  const size_t M = plan_targets->N;
  const double *dg = nullptr, *da, *dt, *du;
  HOLO_TRY(stage_in(ctx, ctx->w, g, 2 * M, dev, &dg));
  HOLO_TRY(stage_in(ctx, ctx->amp, amp, N, dev, &da));
  HOLO_TRY(stage_in(ctx, ctx->phase, theta, N, dev, &dt));
  HOLO_TRY(stage_in(ctx, ctx->u, u, 3 * N, dev, &du));
  double *dga, *dgt;
  HOLO_TRY(out_buf(ctx, ctx->pos, grad_amp, N, dev, &dga));
  HOLO_TRY(out_buf(ctx, ctx->ph, grad_theta, N, dev, &dgt));
  HOLO_TRY(ensure(ctx->b, 2 * N * sizeof(double), ctx->device));
  double* db = static_cast<double*>(ctx->b.ptr);
  const bool obl = flags & HOLO_OBLIQUITY;
  if (M == 0) {
    HOLO_HIP(holo::launch_zero(db, 2 * N, ctx->stream));
  } else {
    // b_n: the forward targets carry g and radiate back with the sign flipped
    // (the obliquity factor |dz| / d is symmetric in the pair)
    HOLO_TRY(plan_run(ctx, plan_targets, precision, -sign, obl, lambda, holo::kCartesian, dg,
                      nullptr, du, N, db));
  }
  HOLO_TRY(report_error(ctx, precision == HOLO_FP64 || M == 0, du, N, db, err_rel_l2, dev,
                        [&](const ErrScratch& e, size_t count) {
                          return plan_fp64_at(ctx, plan_targets, -sign, obl, lambda,
                                              holo::kCartesian, dg, nullptr, e, count);
                        }));
  HOLO_HIP(holo::launch_adjoint_epilogue(reinterpret_cast<const double2*>(db), da, dt, N, dga,
                                         dgt, ctx->stream));
  if (!dev) {
    HOLO_HIP(hipMemcpyAsync(grad_amp, dga, N * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    HOLO_HIP(hipMemcpyAsync(grad_theta, dgt, N * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    HOLO_HIP(hipStreamSynchronize(ctx->stream));
  }
  return HOLO_OK;
}

holo_status holo_plan_layout(const double* u, size_t N, unsigned tile, int32_t* perm,
                             size_t perm_capacity, size_t* n_tiles, double* max_offset,
                             int* planar) {
  if (N > kMaxCount) return fail(HOLO_EINVAL, "N must be below 2^31");
  if (N && !u) return fail(HOLO_EINVAL, "null source positions");
  const unsigned T = tile ? tile : 64;
  if (T != 8 && T != 16 && T != 32 && T != 64)
    return fail(HOLO_EINVAL, "tile must be 8, 16, 32 or 64");
  HostTiles t;
  build_tiles(u, N, T, t);
  if (n_tiles) *n_tiles = t.centre.size();
  if (max_offset) *max_offset = t.max_offset;
  if (planar) *planar = t.planar ? 1 : 0;
  if (perm) {
    if (perm_capacity < t.perm.size()) return fail(HOLO_EINVAL, "perm buffer too small");
    std::copy(t.perm.begin(), t.perm.end(), perm);
  }
  return HOLO_OK;
}

holo_status holo_lattice_layout(const double* u, size_t N, size_t* nx, size_t* ny,
                                size_t* tail, int* tiles16) {
  if (N > kMaxCount) return fail(HOLO_EINVAL, "N must be below 2^31");
  if (N && !u) return fail(HOLO_EINVAL, "null source positions");
  HostGrid g;
  size_t t = 0;
  bool t16 = false;
  if (N && build_grid(u, N, g)) {
    t16 = !g.centre16.empty();
  } else if (N && build_grid16_padded(u, N, g)) {
    t = g.tail.size();
    t16 = true;
  } else {
    g = HostGrid();
  }
  if (nx) *nx = g.nx;
  if (ny) *ny = g.ny;
  if (tail) *tail = t;
  if (tiles16) *tiles16 = t16 ? 1 : 0;
  return HOLO_OK;
}

void holo_host_sincos(double x, double* s, double* c) {
  double a, b;
  holo::hm_sincos(x, a, b);
  if (s) *s = a;
  if (c) *c = b;
}

void holo_host_sqrt_inv(double s, double* d, double* inv) {
  double a, b;
  holo::hm_sqrt_inv(s, a, b);
  if (d) *d = a;
  if (inv) *inv = b;
}

holo_status holo_read_dat(const char* path, double* buf, size_t capacity, size_t* count) {
  if (!path || !count) return fail(HOLO_EINVAL, "null path or count");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(HOLO_EIO, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (size <= 0 || size % (long)sizeof(double)) {
    std::fclose(f);
    return fail(HOLO_EIO, std::string("empty or ragged .dat file ") + path);
  }
  const size_t n = (size_t)size / sizeof(double);
  *count = n;
  if (buf) {
    if (capacity < n) {
      std::fclose(f);
      return fail(HOLO_EINVAL, "buffer too small for " + std::string(path));
    }
    if (std::fread(buf, sizeof(double), n, f) != n) {
      std::fclose(f);
      return fail(HOLO_EIO, std::string("short read of ") + path);
    }
  }
  std::fclose(f);
  return HOLO_OK;
}

holo_status holo_write_dat(const char* path, const double* buf, size_t count) {
  if (!path || (count && !buf)) return fail(HOLO_EINVAL, "null path or buffer");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(HOLO_EIO, std::string("cannot create ") + path);
  const size_t w = count ? std::fwrite(buf, sizeof(double), count, f) : 0;
  const int rc = std::fclose(f);
  if (w != count || rc != 0) return fail(HOLO_EIO, std::string("write failed: ") + path);
  return HOLO_OK;
}

}  // extern "C"
