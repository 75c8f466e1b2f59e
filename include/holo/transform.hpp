// holo/transform.hpp -- C++ drop-in for the reference's host templates, on
// top of the C ABI in holo.h (libholo.so, gfx950).
//
//   transform_full<Direction, Algorithm, shared_memory>(x, u, v, Geometry)
//       cuda/transform.cu:147-340
//   time_transform<Direction, add_constant_wave, add_reference_wave>(...)
//       cuda/transform.cu:346-406
//   geometry(...) / derive_secondary_geometry(...)
//       cuda/init.h:42-84
//
// Semantics kept from the reference:
//   * the returned field has batch_size.y * n_batches.y entries; entries past
//     n.y are the field at padded targets, which sit at the origin
//     (cuda/transform.cu:22-41 zero-fills the last batch);
//   * time_transform divides by the maximum |y| over all those entries and
//     skips the division when it is below 1e-6 (cuda/transform.cu:106-139);
//   * the first source and first target may not share a z offset
//     (cuda/transform.cu:193 asserts u[2] != v[2]);
//   * Algorithm, shared_memory and the block/grid/thread/stream hints of the
//     Geometry select nothing: every algorithm computes the same sum, and the
//     library chooses its own launch (results differ only in summation order).
// Differences: failures throw holo::Error instead of exiting; the reference
// wave is computed from its single source (the reference reuses the object
// plane's batch geometry for it, cuda/transform.cu:388, and so reads past
// its one-element source array).
#pragma once

#include <cmath>
#include <cstdio>
#include <ctime>
#include <stdexcept>
#include <string>
#include <vector>

#include "../holo.h"

#ifndef HOLO_LAMBDA
#define HOLO_LAMBDA 0.65e-6  // cuda/macros.h:23-25 LAMBDA
#endif
#ifndef HOLO_DISTANCE_REFERENCE_WAVE
#define HOLO_DISTANCE_REFERENCE_WAVE 0.25  // cuda/macros.h:26-28
#endif

namespace holo {

constexpr int kFlopPerPoint = 29;  // cuda/macros.h:71-80 FLOP_PER_POINT

enum class Direction { Forwards, Backwards };  // cuda/superposition.cu:19
enum class Algorithm { Naive, Reduced };       // cuda/main.h:7

struct Polar { double amp, phase; };           // cuda/main.h:18
struct Wave { double x, y; };                  // cuDoubleComplex (cuda/macros.h:40)
struct dim2 { size_t x, y; };                  // cuda/main.h:25
struct udim3 { unsigned x = 1, y = 1, z = 1; };
template <typename T> struct Cartesian { T x, y, z; };

struct Plane {                                  // cuda/main.h:28-33
  double width;
  Cartesian<double> offset;
  double aspect_ratio;
  bool randomize;
};

struct Geometry {                               // cuda/main.h:74-86
  int algorithm = 2;
  udim3 blockDim, gridDim, gridSize;
  dim2 n{0, 0}, batch_size{0, 0}, n_batches{0, 0}, thread_size{1, 1};
  size_t n_streams = 1;
};

class Error : public std::runtime_error {
 public:
  Error(holo_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  holo_status status;
};

inline void check(holo_status s, const char* where) {
  if (s != HOLO_OK)
    throw Error(s, std::string(where) + ": " + holo_status_string(s) + ": " + holo_last_error());
}

// One context per host thread on device 0, created on first use.
inline holo_ctx* context() {
  struct Holder {
    holo_ctx* ctx = nullptr;
    ~Holder() {
      if (ctx) holo_destroy(ctx);
    }
  };
  thread_local Holder h;
  if (!h.ctx) check(holo_create(&h.ctx, 0, nullptr), "holo_create");
  return h.ctx;
}

// cuda/init.h:42-60
inline void derive_secondary_geometry(Geometry& p) {
  p.gridSize = {p.blockDim.x * p.gridDim.x, p.blockDim.y * p.gridDim.y, 1};
  p.batch_size = {std::min(p.n.x, p.thread_size.x * p.gridSize.x),
                  std::min(p.n.y, p.thread_size.y * p.gridSize.y)};
  if (p.batch_size.x == 0) p.batch_size.x = 1;
  if (p.batch_size.y == 0) p.batch_size.y = 1;
  p.n_batches = {(p.n.x + p.batch_size.x - 1) / p.batch_size.x,
                 (p.n.y + p.batch_size.y - 1) / p.batch_size.y};
}

// cuda/init.h:73-84, with the defaults of cuda/input.h:68-76.
inline Geometry geometry(size_t n_x, size_t n_y, int algorithm = 2,
                         dim2 thread_size = {256, 128}, udim3 blockDim = {32, 32, 1},
                         udim3 gridDim = {8, 8, 1}, size_t n_streams = 8) {
  Geometry p;
  p.algorithm = algorithm;
  p.blockDim = blockDim;
  p.gridDim = gridDim;
  p.n = {n_x, n_y};
  p.thread_size = thread_size;
  p.n_streams = n_streams;
  derive_secondary_geometry(p);
  return p;
}

inline int sign_of(Direction d) { return d == Direction::Forwards ? HOLO_FORWARDS : HOLO_BACKWARDS; }

// Wavelength and precision used by transform_full / time_transform (the
// reference fixes LAMBDA at compile time, cuda/macros.h:23-25).
struct Settings {
  double lambda = HOLO_LAMBDA;
  holo_precision precision = HOLO_FP64;
  // Cosine (obliquity) factor |dz| / d on every pair, as in the MATLAB
  // light-field model (matlab/holofunctions.m:226-227); the CUDA path has none.
  bool obliquity = false;
  // Error report of the last transform (holo.h): relative L2 distance of an
  // HOLO_FP32_SPLIT field from the FP64 formulation, estimated from 256
  // sampled targets; 0 after an HOLO_FP64 transform.
  double last_rel_l2 = 0.0;
};
inline Settings& settings() {
  static Settings s;
  return s;
}

// Field at `M_out` targets: v is used up to its length and the rest are the
// origin, as the reference's zero-filled last batch.
inline std::vector<Wave> superpose(Direction dir, const std::vector<Polar>& x,
                                   const std::vector<double>& u, const std::vector<double>& v,
                                   size_t M_out, double lambda = settings().lambda,
                                   holo_precision precision = settings().precision) {
  const size_t N = x.size();
  if (u.size() < 3 * N) throw Error(HOLO_EINVAL, "u holds fewer than 3N coordinates");
  std::vector<double> amp(N), phase(N);
  for (size_t i = 0; i < N; ++i) {
    amp[i] = x[i].amp;
    phase[i] = x[i].phase;
  }
  std::vector<double> vv(3 * M_out, 0.0);
  std::copy(v.begin(), v.begin() + std::min(v.size(), vv.size()), vv.begin());
  std::vector<Wave> y(M_out);
  check(holo_transform_full(context(), precision, sign_of(dir), lambda, amp.data(), phase.data(),
                            u.data(), N, vv.data(), M_out, reinterpret_cast<double*>(y.data()),
                            &settings().last_rel_l2,
                            settings().obliquity ? HOLO_OBLIQUITY : 0u),
        "holo_transform_full");
  return y;
}

// cuda/transform.cu:147-340
template <Direction direction, Algorithm algorithm = Algorithm::Naive, bool shared_memory = false>
inline std::vector<Wave> transform_full(const std::vector<Polar>& x, const std::vector<double>& u,
                                        const std::vector<double>& v, const Geometry& p) {
  static_assert(!(algorithm == Algorithm::Naive && shared_memory),
                "Naive has no shared-memory variant (cuda/transform.cu:157)");
  if (u.size() >= 3 && v.size() >= 3 && u[2] == v[2])
    throw Error(HOLO_EINVAL, "u[2] == v[2]: first source and target share a z offset "
                             "(cuda/transform.cu:193)");
  return superpose(direction, x, u, v, p.batch_size.y * p.n_batches.y);
}

// cuda/transform.cu:106-139 (without the log-normalisation branch)
template <bool add_constant = false>
inline void normalize_amp(std::vector<Wave>& c, double to = 1.) {
  double max_amp = 0;
  for (const auto& w : c) max_amp = std::fmax(max_amp, std::hypot(w.x, w.y));
  if (max_amp < 1e-6) {
    std::printf("WARNING, max_amp << 1\n");
    return;
  }
  max_amp /= to;
  if (add_constant) max_amp *= 2.;
  for (auto& w : c) {
    w.x = w.x / max_amp + (add_constant ? 0.5 : 0.0);
    w.y = w.y / max_amp;
  }
}

inline double diff(const timespec& t0, const timespec& t1) {
  return (double)(t1.tv_sec - t0.tv_sec) + (double)(t1.tv_nsec - t0.tv_nsec) * 1e-9;
}

inline double flops(double runtime, size_t n, size_t m) {  // cuda/util.h:20-31
  return n * m * (double)kFlopPerPoint / runtime;
}

// cuda/transform.cu:346-406
template <Direction direction, bool add_constant_wave = false, bool add_reference_wave = false>
inline std::vector<Polar> time_transform(const std::vector<Polar>& x, const std::vector<double>& u,
                                         const std::vector<double>& v, const Geometry& p,
                                         const Plane& /*x_plane*/, const Plane& /*y_plane*/,
                                         timespec* t1, timespec* t2, double* dt,
                                         bool verbose = false) {
  static_assert(!add_constant_wave, "add_constant_wave is unsupported (cuda/transform.cu:378)");
  clock_gettime(CLOCK_MONOTONIC, t1);
  if (p.algorithm < 1 || p.algorithm > 3) throw Error(HOLO_EINVAL, "algorithm is incorrect");
  std::vector<Wave> y = transform_full<direction, Algorithm::Reduced>(x, u, v, p);
  const double rel_l2 = settings().last_rel_l2;  // of this field, not the reference wave's
  normalize_amp<false>(y, 1.0);
  if (add_reference_wave) {
    // single far-away source behind the v plane, unit amplitude, phase 0
    const double z_offset = v[2] - HOLO_DISTANCE_REFERENCE_WAVE;
    auto ref = superpose(Direction::Forwards, {{1.0, 0.0}}, {0.0, 0.0, z_offset}, v, y.size());
    normalize_amp<false>(ref, 1.0);
    for (size_t i = 0; i < y.size(); ++i) {
      y[i].x += ref[i].x;
      y[i].y += ref[i].y;
    }
    for (size_t i = 0; i < p.n.y && i < y.size(); ++i) {
      const double a = std::hypot(y[i].x, y[i].y), phi = std::atan2(ref[i].y, ref[i].x);
      y[i] = {a * std::cos(phi), a * std::sin(phi)};
    }
  }
  clock_gettime(CLOCK_MONOTONIC, t2);
  *dt = diff(*t1, *t2);
  settings().last_rel_l2 = rel_l2;
  if (verbose) {
    std::printf("TFLOPS:   \t%0.5f \t (%i FLOP_PER_POINT) \t E[runtime]: %f s\n",
                flops(*dt, x.size(), y.size()) * 1e-12, kFlopPerPoint, *dt);
    if (settings().precision == HOLO_FP32_SPLIT)
      std::printf("FP32-split rel. L2 vs FP64 (256 targets): %.3e\n", rel_l2);
  }
  std::vector<Polar> out(std::min(p.n.y, y.size()));
  for (size_t i = 0; i < out.size(); ++i)
    out[i] = {std::hypot(y[i].x, y[i].y), std::atan2(y[i].y, y[i].x)};
  return out;
}

}  // namespace holo
