// The C++ drop-in (include/holo/transform.hpp) used as cuda/main.cu uses the
// reference's templates (cuda/main.cu:156-158, 186-188): a small seeded problem
// through transform_full<Backwards, Reduced, shared_memory> and
// time_transform<Backwards>, and the error the reference asserts on
// (cuda/transform.cu:193).  Writes into argv[1]: the inputs (x_amp, x_phase,
// u, v), the unnormalised field y (interleaved, padded length) and the
// normalised polar output (p_amp, p_phase); prints "caught EINVAL" for the
// error case.  tests/test_cli.py compares the files with the oracle.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "holo/transform.hpp"

namespace {

uint64_t g_state = 0x9e3779b97f4a7c15ull;
double uniform() {  // splitmix64 -> [0, 1)
  uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1p-53;
}

void write(const std::string& dir, const char* name, const double* p, size_t n) {
  holo::check(holo_write_dat((dir + "/" + name).c_str(), p, n), name);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: transform_dropin OUTDIR\n");
    return 2;
  }
  const std::string dir = argv[1];
  const size_t nx = 24, N = nx * nx, my = 40, M = my * my;
  std::vector<holo::Polar> x(N);
  std::vector<double> u(3 * N), v(3 * M), amp(N), phase(N);
  for (size_t i = 0; i < N; ++i) {
    x[i] = {uniform(), 2 * M_PI * uniform()};
    amp[i] = x[i].amp;
    phase[i] = x[i].phase;
    u[3 * i] = ((double)(i / nx) - (nx - 1) / 2.0) * 10e-6;
    u[3 * i + 1] = ((double)(i % nx) - (nx - 1) / 2.0) * 10e-6;
    u[3 * i + 2] = 0.0;
  }
  for (size_t m = 0; m < M; ++m) {
    v[3 * m] = ((double)(m / my) - (my - 1) / 2.0) * 7.56e-6;
    v[3 * m + 1] = ((double)(m % my) - (my - 1) / 2.0) * 7.56e-6;
    v[3 * m + 2] = 0.35;
  }
  // 4 targets per thread (the reference's -T hint) make batches of 1024
  // targets, so the output is padded to 2048 with targets at the origin
  const holo::Geometry p = holo::geometry(N, M, 2, {256, 4});
  const auto y = holo::transform_full<holo::Direction::Backwards, holo::Algorithm::Reduced, true>(
      x, u, v, p);
  const holo::Plane plane{1e-3, {0., 0., 0.}, 1., false};
  timespec t1, t2;
  double dt = 0.0;
  const auto pol = holo::time_transform<holo::Direction::Backwards>(x, u, v, p, plane, plane, &t1,
                                                                    &t2, &dt);
  std::vector<double> pa(pol.size()), pp(pol.size());
  for (size_t i = 0; i < pol.size(); ++i) {
    pa[i] = pol[i].amp;
    pp[i] = pol[i].phase;
  }
  write(dir, "x_amp.dat", amp.data(), N);
  write(dir, "x_phase.dat", phase.data(), N);
  write(dir, "u.dat", u.data(), u.size());
  write(dir, "v.dat", v.data(), v.size());
  write(dir, "y.dat", reinterpret_cast<const double*>(y.data()), 2 * y.size());
  write(dir, "p_amp.dat", pa.data(), pa.size());
  write(dir, "p_phase.dat", pp.data(), pp.size());
  std::vector<double> v0 = v;
  v0[2] = 0.0;  // first target in the first source's plane
  try {
    holo::transform_full<holo::Direction::Backwards>(x, u, v0, p);
    std::printf("no error\n");
  } catch (const holo::Error& e) {
    std::printf("caught %s\n", e.status == HOLO_EINVAL ? "EINVAL" : "other");
  }
  return 0;
}
