// holo -- command-line driver with the flags and files of the reference's
// `holo` executable (cuda/main.cu, cuda/input.h), running the superposition
// on an MI355X through libholo.so.
//
// Modes (cuda/main.cu:34-227):
//   -f DIR [-F]   read DIR/x0_amp.dat, x0_phase.dat, u0.dat, v0.dat (and w0.dat
//                 with -F), transform x -> v backwards, write y0_amp.dat,
//                 y0_phase.dat, v0.dat and out.json into DIR.  With -F the
//                 reference computes the x -> v transform forwards and writes
//                 it as z0_0 / w0_0 (cuda/main.cu:218-221); that is kept.
//   (no -f)       generate X object planes (circle of -x points, or the line,
//                 log-line or dotted circle HOLO_SHAPE selects) and the
//                 projector / projection lattices, run x -> y (backwards) and
//                 y -> z (forwards) per plane pair, write to ../tmp/.
// The target count is -y (not the size of v0.dat); missing targets are the
// origin, as the reference's zero-filled batches.  Geometry flags (-s -t -T
// -b -B -g -G) only set the padded output length used for normalisation.
// HOLO_PROJECT_PHASE=0 in the environment selects the amplitude-only
// projector with a reference wave (the reference's PROJECT_PHASE=0 build).
// Every error prints a message and exits with status 1 (the reference
// aborts on a failed assert).
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "holo/transform.hpp"

using holo::Cartesian;
using holo::Direction;
using holo::Geometry;
using holo::Plane;
using holo::Polar;

namespace {

constexpr size_t kNSqrt = 512;  // cuda/macros.h:8 N_sqrt

template <typename T> struct Range { T min, max; };
template <typename T> struct Setup { T obj, projector, projection; };

struct Params {  // cuda/main.h:40-64, defaults of cuda/input.h:44-79
  Setup<size_t> n_planes{1, 1, 1};
  Setup<size_t> n_per_plane{1, kNSqrt * kNSqrt, kNSqrt * kNSqrt};
  Setup<double> aspect_ratio{1., 1., 1.};
  Cartesian<Range<double>> obj_offset{{0.0, 0.0}, {0.0, 0.0}, {0.35, 0.25}};
  Range<double> obj_width{1e-3, 1e-3};
  Range<double> projector_width{kNSqrt * 7e-6, kNSqrt * 7e-6};
  Range<double> projection_width{1e-3 * 1.2, 1e-3 * 1.2};
  Range<double> projection_z_offset{0., 0.};
  int algorithm = 2;
  bool quadrant_projection = false, randomize = false;
  unsigned n_streams = 8;
  holo::dim2 thread_size{256, 128};
  holo::udim3 blockDim{32, 32, 1}, gridDim{8, 8, 1};
  std::string files;
  bool projection = false;
};

[[noreturn]] void die(const std::string& msg) {
  fprintf(stderr, "holo: %s\n", msg.c_str());
  exit(1);
}

void require(bool ok, const char* what) {
  if (!ok) die(std::string("invalid arguments: ") + what);
}

void show_help(const char* p) {  // cuda/input.h:7-39
  printf("Usage %s [OPTION]...\n"
         "Note: for each attribute, if .max is not provided it defaults to the corresponding .min.\n"
         "\t -h       \tShow help. \n"
         "\n\tUse external data (optional)\n"
         "\t -f {dir} \t Read input files from `dir`: `x0_amp.dat, x0_phase.dat, u0.dat, v0.dat`."
         " These files are raw arrays of little-endian float64.\n"
         "\t-F        \t Also read w0.dat and write the projection (as z0_0, w0_0).\n"
         "\n\tPerformance parameters (accepted; the library picks its own launch)\n"
         "\t-s        \tNumber of streams (default: 8).\n"
         "\t-t, -T    \tSource, target datapoints per thread (default 256, 128).\n"
         "\t-b, -B    \tblockDim .x, .y (default 32, 32).\n"
         "\t-g, -G    \tgridDim .x, .y (default 8, 8).\n"
         "\t-p        \tAlgorithm 1, 2 or 3 (default: 2); all compute the same sum.\n"
         "\n\tGenerate data (used if `-f {dir}` is not supplied).\n"
         "\t-X, -Z    \tNumber of input, projection planes.\n"
         "\t-x, -y, -z\tNumber of input, projector, projection datapoints per plane.\n"
         "\t-a, -A    \tAspect ratio of projector, projection.\n"
         "\t-u, -U    \t.min, .max offset in x-dimension for input objects.\n"
         "\t-v, -V    \t.min, .max offset in y-dimension for input objects.\n"
         "\t-w, -W    \t.min, .max offset in z-dimension for input objects.\n"
         "\t-o, -O    \t.min, .max object width.\n"
         "\t-l, -L    \t.min, .max projector width.\n"
         "\t-n, -N    \t.min, .max projection width.\n"
         "\t-m, -M    \t.min, .max projection offset in z-dimension.\n"
         "\t-q        \tLimit projection to a single quadrant.\n"
         "\t-r        \tRandomize projector pixel positions.\n"
         "\n\tEnvironment\n"
         "\t HOLO_LAMBDA \t Wavelength in m (default 0.65e-6)\n"
         "\t HOLO_PROJECT_PHASE \t 0: amplitude-only projector with reference wave (default 1)\n"
         "\t HOLO_PRECISION \t fp64 (default) or fp32split\n"
         "\t HOLO_OBLIQUITY \t 1: weight every pair by |dz|/d (default 0)\n"
         "\t HOLO_SHAPE \t generated object: circle (default), line, logline, dottedcircle\n"
         "\n", p);
}

Params read_args(int argc, char** argv) {  // cuda/input.h:43-151
  Params p;
  int ch;
  while ((ch = getopt(argc, argv, "X:Z:x:y:z:a:A:u:U:v:V:w:W:o:O:l:L:n:N:m:M:p:qrs:t:T:b:B:g:G:f:Fh")) != -1) {
    switch (ch) {
      case 'X': p.n_planes.obj = strtol(optarg, 0, 10); break;
      case 'Z': p.n_planes.projection = strtol(optarg, 0, 10); break;
      case 'x': p.n_per_plane.obj = strtol(optarg, 0, 10); break;
      case 'y': p.n_per_plane.projector = strtol(optarg, 0, 10); break;
      case 'z': p.n_per_plane.projection = strtol(optarg, 0, 10); break;
      case 'a': p.aspect_ratio.projector = strtod(optarg, 0); break;
      case 'A': p.aspect_ratio.projection = strtod(optarg, 0); break;
      case 'u': p.obj_offset.x.min = p.obj_offset.x.max = strtod(optarg, 0); break;
      case 'U': p.obj_offset.x.max = strtod(optarg, 0); break;
      case 'v': p.obj_offset.y.min = p.obj_offset.y.max = strtod(optarg, 0); break;
      case 'V': p.obj_offset.y.max = strtod(optarg, 0); break;
      case 'w': p.obj_offset.z.min = p.obj_offset.z.max = strtod(optarg, 0); break;
      case 'W': p.obj_offset.z.max = strtod(optarg, 0); break;
      case 'o': p.obj_width.min = p.obj_width.max = strtod(optarg, 0); break;
      case 'O': p.obj_width.max = strtod(optarg, 0); break;
      case 'l': p.projector_width.min = p.projector_width.max = strtod(optarg, 0); break;
      case 'L': p.projector_width.max = strtod(optarg, 0); break;
      case 'n': p.projection_width.min = p.projection_width.max = strtod(optarg, 0); break;
      case 'N': p.projection_width.max = strtod(optarg, 0); break;
      case 'm': p.projection_z_offset.min = p.projection_z_offset.max = strtod(optarg, 0); break;
      case 'M': p.projection_z_offset.max = strtod(optarg, 0); break;
      case 'p': p.algorithm = strtol(optarg, 0, 10); break;
      case 'q': p.quadrant_projection = true; break;
      case 'r': p.randomize = true; break;
      case 's': p.n_streams = strtol(optarg, 0, 10); break;
      case 't': p.thread_size.x = strtol(optarg, 0, 10); break;
      case 'T': p.thread_size.y = strtol(optarg, 0, 10); break;
      case 'b': p.blockDim.x = strtol(optarg, 0, 10); break;
      case 'B': p.blockDim.y = strtol(optarg, 0, 10); break;
      case 'g': p.gridDim.x = strtol(optarg, 0, 10); break;
      case 'G': p.gridDim.y = strtol(optarg, 0, 10); break;
      case 'f': p.files = std::string{optarg} + "/"; break;
      case 'F': p.projection = true; break;
      case 'h':
      default: show_help(argv[0]); exit(0);
    }
  }
  require(0 < p.algorithm && p.algorithm <= 3, "algorithm must be 1, 2 or 3");
  require(p.n_planes.obj >= 1, "-X must be >= 1");
  require(p.n_planes.projection <= 10000, "-Z must be <= 10000");
  require(p.n_per_plane.projector > 0, "-y must be > 0");
  if (p.n_planes.projection)
    require(p.n_per_plane.projection > 0, "-z must be > 0");
  else
    p.n_per_plane.projection = 1;
  require(p.blockDim.x * p.blockDim.y <= 1024, "blockDim.x * blockDim.y must be <= 1024");
  require(p.thread_size.x > 0 && p.thread_size.y > 0 && p.n_streams > 0 && p.gridDim.x > 0 &&
              p.gridDim.y > 0 && p.blockDim.x > 0 && p.blockDim.y > 0,
          "geometry hints must be positive");
  const double nonzero = 1e-6;
  if (p.obj_offset.z.min > 0.)
    require(p.obj_offset.z.min > nonzero && p.obj_offset.z.max > nonzero, "object z offset too small");
  else
    require(-p.obj_offset.z.min > nonzero && -p.obj_offset.z.max > nonzero, "object z offset too small");
  require(std::fabs(p.obj_offset.z.min + p.projection_z_offset.min) > nonzero &&
              std::fabs(p.obj_offset.z.min + p.projection_z_offset.max) > nonzero &&
              std::fabs(p.obj_offset.z.max + p.projection_z_offset.max) > nonzero &&
              std::fabs(p.obj_offset.z.max + p.projection_z_offset.min) > nonzero,
          "projection plane coincides with the projector plane");
  return p;
}

std::vector<double> linspace(size_t len, double min, double max) {  // cuda/algebra.h:53-74
  std::vector<double> v(len, min);
  if (len <= 1) return v;
  const double delta = (max - min) / (len - 1.0);
  for (size_t i = 0; i < len; ++i) v[i] += delta * i;
  return v;
}

double lerp(Range<double> r, double ratio) { return (1 - ratio) * r.min + ratio * r.max; }

double gerp(Range<double> r, double ratio) {  // cuda/algebra.h:140-150
  double a = r.min, b = r.max;
  if (a == b) return a;
  if (a == 0) a = 1e-12;
  if (b == 0) b = 1e-12;
  return std::pow(10, (1 - ratio) * std::log10(a) + ratio * std::log10(b));
}

// cuda/init.h:89-158: centred lattice, flat index i*ny + j; remaining points
// copy point 0.  -r jitters each pixel (the reference draws from cuRAND;
// here std::mt19937_64, so the jitter values differ).
void plane(std::vector<double>& v, const Plane& p) {
  const size_t n = v.size() / 3;
  const size_t x = (size_t)std::sqrt(n * p.aspect_ratio), y = x ? n / x : 0;
  const double height = p.width / p.aspect_ratio, dx = p.width / (double)x,
               dy = height / (double)y, x_half = (p.width - dx) / 2., y_half = (height - dy) / 2.;
  static uint64_t seed = 1234;
  std::mt19937_64 rng(++seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const double margin = p.randomize ? 0.05 : 0.0;
  const double xr = dx - 0.5 * margin * dx, yr = dy - 0.5 * margin * dy;
  for (size_t i = 0; i < x; ++i)
    for (size_t j = 0; j < y; ++j) {
      const size_t k = 3 * (i * y + j);
      v[k] = i * dx - x_half + p.offset.x;
      v[k + 1] = j * dy - y_half + p.offset.y;
      v[k + 2] = p.offset.z;
      if (p.randomize) {
        v[k] += xr * (U(rng) - 0.5);
        v[k + 1] += yr * (U(rng) - 0.5);
      }
    }
  for (size_t i = x * y; i < n; ++i)
    for (int d = 0; d < 3; ++d) v[3 * i + d] = n ? v[d] : 0.0;
}

// cuda/main.h:9; cuda/main.cu:55 generates a Circle, HOLO_SHAPE another.
enum class Shape { Line, LogLine, Circle, DottedCircle };
Shape g_shape = Shape::Circle;

// cuda/algebra.h:77-91: len values from a to b, evenly spaced in log10
std::vector<double> geomspace(size_t len, double a, double b) {
  std::vector<double> v = linspace(len, std::log10(a), std::log10(b));
  for (double& x : v) x = std::pow(10., x);
  return v;
}

// cuda/init.h:160-234: n points of the shape at z = offset.z, then shifted by
// (offset.x, offset.y) when either is nonzero.  Line: x from -w/2 to w/2;
// LogLine: x = geomspace(1, 1 + w) - 1 + w/2; Circle: n points on the circle
// of diameter w, starting at angle 0.1125 + 2 pi modulate; DottedCircle: point
// 0 at the centre, n - 1 points on the circle.  y is 0 on the lines.  (The
// reference leaves point 0 of a DottedCircle as its reused vector held it: the
// centre on the first plane.)
void sparse_plane(std::vector<double>& u, Shape shape, double width,
                  const Cartesian<double>& offset, double modulate) {
  const size_t n = u.size() / 3;
  for (size_t i = 0; i < n; ++i) {
    u[3 * i] = u[3 * i + 1] = 0.0;
    u[3 * i + 2] = offset.z;
  }
  if (n == 0) return;
  if (n == 1) {
    u[0] = offset.x;
    u[1] = offset.y;
    return;
  }
  const double radius = width / 2.0, arbitrary_offset = 0.1125 + modulate * 2 * M_PI;
  switch (shape) {
    case Shape::Line: {
      const double du = width / (double)(n - 1), half_width = width / 2.0;
      for (size_t i = 0; i < n; ++i) u[3 * i] = i * du - half_width;
      break;
    }
    case Shape::LogLine: {
      const std::vector<double> x = geomspace(n, 1, 1 + width);
      for (size_t i = 0; i < n; ++i) u[3 * i] = x[i] - 1 + width / 2.;
      break;
    }
    case Shape::Circle: {
      const double d_phase = 2 * M_PI / (double)n;
      for (size_t i = 0; i < n; ++i) {
        u[3 * i] = std::sin(i * d_phase + arbitrary_offset) * radius;
        u[3 * i + 1] = std::cos(i * d_phase + arbitrary_offset) * radius;
      }
      break;
    }
    case Shape::DottedCircle: {
      const double d_phase = 2 * M_PI / (double)(n - 1);
      for (size_t i = 1; i < n; ++i) {
        u[3 * i] = std::sin(i * d_phase + arbitrary_offset) * radius;
        u[3 * i + 1] = std::cos(i * d_phase + arbitrary_offset) * radius;
      }
      break;
    }
  }
  if (offset.x != 0 || offset.y != 0)
    for (size_t i = 0; i < n; ++i) {
      u[3 * i] += offset.x;
      u[3 * i + 1] += offset.y;
    }
}

std::vector<double> read_dat(const std::string& path) {  // cuda/util.h:248-283
  size_t n = 0;
  if (holo_read_dat(path.c_str(), nullptr, 0, &n) != HOLO_OK) die(holo_last_error());
  std::vector<double> buf(n);
  if (holo_read_dat(path.c_str(), buf.data(), n, &n) != HOLO_OK) die(holo_last_error());
  return buf;
}

// cuda/util.h:358-444: out.json line + <k1>_amp.dat, <k1>_phase.dat, <k2>.dat.  A
// field computed in FP32-split (computed = true) also records its error report
// as "rel_l2_fp64" (holo.h); FP64 runs write the reference's keys only.
void write_arrays(const std::vector<Polar>& x, const std::vector<double>& u, const std::string& k1,
                  const std::string& k2, const Plane& p, double dt, double fl, const std::string& dir,
                  bool computed = false) {
  static bool overwrite = true;
  std::ofstream out(dir + "out.json", overwrite ? std::ofstream::binary : std::ofstream::app);
  if (!out) die("cannot write " + dir + "out.json");
  double amp = 0, phase = 0;
  for (const auto& e : x) {
    amp += e.amp;
    phase += e.phase;
  }
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  auto q = [](const std::string& s) { return "\"" + s + "\""; };
  out << "{" << q("phasor") << ": " << q(k1) << ", " << q("amp_sum") << ": " << amp << ", "
      << q("phase_sum") << ": " << phase << ", " << q("pos") << ": " << q(k2) << ", " << q("len")
      << ": " << x.size() << ", " << q("precision") << ": " << 8 << ", " << q("dims") << ": " << 3
      << ", " << q("x_offset") << ": " << p.offset.x << ", " << q("y_offset") << ": " << p.offset.y
      << ", " << q("z_offset") << ": " << p.offset.z << ", " << q("width") << ": " << p.width
      << ", " << q("randomized") << ": " << p.randomize << ", " << q("aspect_ratio") << ": "
      << p.aspect_ratio << ", " << q("timestamp") << ": " << t.tv_nsec << ", " << q("runtime")
      << ": " << dt << ", " << q("flops") << ": " << fl;
  if (computed && holo::settings().precision == HOLO_FP32_SPLIT)
    out << ", " << q("rel_l2_fp64") << ": " << holo::settings().last_rel_l2;
  out << "}\n";
  out.close();
  if (!std::isfinite(amp) || !std::isfinite(phase)) die("non-finite result in " + k1);
  std::vector<double> a(x.size()), ph(x.size());
  for (size_t i = 0; i < x.size(); ++i) {
    a[i] = x[i].amp;
    ph[i] = x[i].phase;
  }
  if (holo_write_dat((dir + k1 + "_amp.dat").c_str(), a.data(), a.size()) != HOLO_OK ||
      holo_write_dat((dir + k1 + "_phase.dat").c_str(), ph.data(), ph.size()) != HOLO_OK ||
      holo_write_dat((dir + k2 + ".dat").c_str(), u.data(), u.size()) != HOLO_OK)
    die(holo_last_error());
  overwrite = false;
}

template <Direction dir, bool add_ref>
std::vector<Polar> run(const std::vector<Polar>& x, const std::vector<double>& u,
                       const std::vector<double>& v, const Geometry& g, const Plane& a,
                       const Plane& b, double* dt, bool verbose) {
  timespec t1, t2;
  try {
    return holo::time_transform<dir, false, add_ref>(x, u, v, g, a, b, &t1, &t2, dt, verbose);
  } catch (const holo::Error& e) {
    die(e.what());
  }
}

template <bool add_ref>
int body(Params& params) {
  const bool file_mode = !params.files.empty();
  std::vector<Polar> x;
  std::vector<double> u, v, w;
  auto& npp = params.n_per_plane;
  if (!file_mode) {
    u.assign(3 * npp.obj, 0.0);
    v.assign(3 * npp.projector, 0.0);
    const auto amp = linspace(npp.obj, 1, 0);  // descending intensity (cuda/main.cu:65-73)
    x.resize(npp.obj);
    for (size_t i = 0; i < x.size(); ++i) x[i] = {amp[i], 0.0};
  } else {
    const auto a = read_dat(params.files + "x0_amp.dat");
    const auto ph = read_dat(params.files + "x0_phase.dat");
    if (a.size() != ph.size()) die("x0_amp.dat and x0_phase.dat differ in length");
    x.resize(a.size());
    for (size_t i = 0; i < x.size(); ++i) x[i] = {a[i], ph[i]};
    u = read_dat(params.files + "u0.dat");
    v = read_dat(params.files + "v0.dat");
    if (params.projection) w = read_dat(params.files + "w0.dat");
    npp.obj = x.size();
    if (x.size() != u.size() / 3) die("x0_amp.dat and u0.dat describe different source counts");
    printf("Number of input datapoints/plane: %zu\n", x.size());
  }
  if (!file_mode || !params.projection) w.assign(3 * npp.projection, 0.0);
  auto geom = [&](size_t nx, size_t ny) {
    return holo::geometry(nx, ny, params.algorithm, params.thread_size, params.blockDim,
                          params.gridDim, params.n_streams);
  };
  const Geometry projector = geom(npp.obj, npp.projector),
                 projection = geom(npp.projector, npp.projection);
  std::vector<double> dt(std::max<size_t>(params.n_planes.projection, 1), 0.0);

  if (!file_mode) {
    for (size_t i = 0; i < params.n_planes.obj; ++i) {
      const double ratio = i == 0 ? 0 : i / ((double)params.n_planes.obj - 1.);
      const Cartesian<double> obj_offset{lerp(params.obj_offset.x, ratio),
                                         lerp(params.obj_offset.y, ratio),
                                         gerp(params.obj_offset.z, ratio)};
      const Plane x_plane{lerp(params.obj_width, ratio), obj_offset, 1., false};
      sparse_plane(u, g_shape, x_plane.width, x_plane.offset, i / (double)params.n_planes.obj);
      const Plane y_plane{lerp(params.projector_width, ratio), {0., 0., 0.},
                          params.aspect_ratio.projector, params.randomize};
      plane(v, y_plane);
      const std::string xs = std::to_string(i);
      write_arrays(x, u, "x" + xs, "u" + xs, x_plane, 0., 0., "../tmp/");
      const auto y = run<Direction::Backwards, add_ref>(x, u, v, projector, x_plane, y_plane,
                                                        &dt[0], false);
      write_arrays(y, v, "y" + xs, "v" + xs, y_plane, dt[0], holo::flops(dt[0], x.size(), y.size()),
                   "../tmp/", true);
      for (size_t j = 0; j < params.n_planes.projection; ++j) {
        const double r2 = j == 0 ? 0 : j / ((double)params.n_planes.projection - 1.);
        const double width = gerp(params.projection_width, r2),
                     height = width / params.aspect_ratio.projection;
        const Cartesian<double> po{params.quadrant_projection ? width / 2. : 0.,
                                   params.quadrant_projection ? height / 2. : 0.,
                                   lerp(params.projection_z_offset, r2)};
        const Plane z_plane{width,
                            {obj_offset.x + po.x, obj_offset.y + po.y, obj_offset.z + po.z},
                            params.aspect_ratio.projection, params.randomize};
        plane(w, z_plane);
        const auto z = run<Direction::Forwards, false>(y, v, w, projection, y_plane, z_plane,
                                                       &dt[j], false);
        const std::string zs = xs + "_" + std::to_string(j);
        write_arrays(z, w, "z" + zs, "w" + zs, z_plane, dt[j],
                     holo::flops(dt[j], y.size(), z.size()), "../tmp/", true);
      }
    }
  } else {
    const Plane x_plane{1, {0., 0., 1.}, 1, false}, y_plane{1, {0., 0., 0.}, 1, false},
        z_plane{1, {0., 0., 1.}, 1, false};
    const auto y = run<Direction::Backwards, add_ref>(x, u, v, projector, x_plane, y_plane, &dt[0],
                                                      true);
    write_arrays(y, v, "y0", "v0", y_plane, dt[0], holo::flops(dt[0], x.size(), y.size()),
                 params.files, true);
    if (params.projection) {
      const auto z = run<Direction::Forwards, false>(x, u, v, projector, x_plane, y_plane, &dt[0],
                                                     true);
      write_arrays(z, w, "z0_0", "w0_0", z_plane, dt[0], holo::flops(dt[0], y.size(), z.size()),
                   params.files, true);
    }
  }
  printf("done\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Params params = read_args(argc, argv);
  if (const char* l = getenv("HOLO_LAMBDA")) {
    char* end = nullptr;
    const double lam = strtod(l, &end);
    if (!(lam > 0.0) || end == l) die("HOLO_LAMBDA must be a positive wavelength in metres");
    holo::settings().lambda = lam;
  }
  if (const char* pr = getenv("HOLO_PRECISION")) {
    const std::string v(pr);
    if (v == "fp64") holo::settings().precision = HOLO_FP64;
    else if (v == "fp32split") holo::settings().precision = HOLO_FP32_SPLIT;
    else die("HOLO_PRECISION must be fp64 or fp32split");
  }
  if (const char* ob = getenv("HOLO_OBLIQUITY")) {
    const std::string v(ob);
    if (v == "1") holo::settings().obliquity = true;
    else if (v != "0") die("HOLO_OBLIQUITY must be 0 or 1");
  }
  if (const char* sh = getenv("HOLO_SHAPE")) {
    const std::string v(sh);
    if (v == "circle") g_shape = Shape::Circle;
    else if (v == "line") g_shape = Shape::Line;
    else if (v == "logline") g_shape = Shape::LogLine;
    else if (v == "dottedcircle") g_shape = Shape::DottedCircle;
    else die("HOLO_SHAPE must be circle, line, logline or dottedcircle");
  }
  const char* pp = getenv("HOLO_PROJECT_PHASE");
  const bool project_phase = !(pp && std::string(pp) == "0");
  return project_phase ? body<false>(params) : body<true>(params);
}
