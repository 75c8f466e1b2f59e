/*
 * holo_oracle.c -- CPU oracle for the Huygens point-source superposition.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * It restates, in plain C with every floating-point operation written out,
 * the FP64 formulation of the reference:
 *
 *   pair      cuda/superposition.cu:35-51   phasor_displacement<Direction>
 *             cuda/kernel.cu:148-166        cos_sin / from_polar
 *             cuda/macros.h:23-31           TWO_PI_OVER_LAMBDA = (2*M_PI)/LAMBDA
 *   sum       cuda/superposition.cu:101-111 per-target sum over all sources
 *             cuda/transform.cu:243-299     batching + row sum (order-free)
 *   polar     cuda/transform.cu:106-139     normalize_amp (divide by max |y|)
 *             cuda/transform.cu:400-405     Polar{|y|, atan2(im, re)}
 *
 * Pinned operation order (the reference leaves it to nvcc/libdevice):
 *   dx = v0-u0; dy = v1-u1; dz = v2-u2
 *   s   = fma(dz, dz, fma(dy, dy, dx*dx))       a*a+b*b+c*c, contracted
 *   d   = sqrt(s)                                correctly rounded
 *   arg = fma(sign*d, K, phi)                    phase +/- distance*K, contracted
 *   r   = a / d                                  correctly rounded
 *   re  = r*cos(arg), im = r*sin(arg)
 * nvcc contracts `x.phase + distance * TWO_PI_OVER_LAMBDA` into one fused
 * multiply-add by default (--fmad=true), so the single rounding of `arg` is
 * the reference's own effective order on its GPU.  Build with
 * -ffp-contract=off so that gcc adds no contraction of its own.
 *
 * Monte Carlo (no counterpart left in the reference snapshot; contract from
 * cuda/kernel.cu:101-106 `randint` and the benchmark harness): target m draws
 * K source indices n_{m,k} uniformly with replacement and the estimate is
 * (N/K) * sum_k pair(n_{m,k}, m).  Indices come from Philox4x32-10
 * (Salmon et al., SC'11) keyed by the 64-bit seed with counter
 * (k/2, m_lo, m_hi, 0x4d43); word pair (0,1) serves even k, (2,3) odd k; the
 * 64-bit word r maps to n = floor(r*N / 2^64) (Lemire's multiply-shift).
 *
 * Pinning status of each function is stated beside it (see tests/).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_TWO_PI (2 * M_PI) /* cuda/macros.h:29 */

/* K = TWO_PI_OVER_LAMBDA (cuda/macros.h:31), evaluated in double. */
double oracle_wavenumber(double lambda) { return ORACLE_TWO_PI / lambda; }

/* One pair: superposition.cu:35-51.  sign = +1 Forwards, -1 Backwards.
 * Also returns d and arg so that tests can check them bit-for-bit.
 * obliquity != 0 multiplies the amplitude by the obliquity factor
 * cos = |v2 - u2| / d of matlab/holofunctions.m:226-227 (`cosfact = Gz ./
 * Gdistances`, which assumes one plane separation Gz; here the per-pair
 * axial distance): r = (a / d) * (|dz| / d).
 * Pinned by: py/test_util.py:63-64 known answer, the exact-rational pure
 * Python restatement in tests/test_oracle.py, mpmath closed forms; the
 * obliquity factor by the closed form a |z| / d^2 and a NumPy restatement of
 * lightfieldcalc. */
void oracle_pair(int sign, double K, double a, double phi,
                 const double *u, const double *v, int obliquity,
                 double *re, double *im, double *d_out, double *arg_out)
{
    const double dx = v[0] - u[0], dy = v[1] - u[1], dz = v[2] - u[2];
    const double s = fma(dz, dz, fma(dy, dy, dx * dx));
    const double d = sqrt(s);
    const double sd = sign > 0 ? d : -d;
    const double arg = fma(sd, K, phi);
    double r = a / d;
    if (obliquity) r *= fabs(dz) / d;
    *re = r * cos(arg);
    *im = r * sin(arg);
    if (d_out) *d_out = d;
    if (arg_out) *arg_out = arg;
}

/* Same pair with a complex source weight w = (w_re, w_im) instead of a
 * polar one: w * exp(i sign d K) / d.  This is the adjoint's pair
 * (conj of the forward kernel when sign is flipped); no reference
 * counterpart.  Pinned by the dense-matrix identity in tests. */
void oracle_pair_cartesian(int sign, double K, double w_re, double w_im,
                           const double *u, const double *v, int obliquity,
                           double *re, double *im)
{
    const double dx = v[0] - u[0], dy = v[1] - u[1], dz = v[2] - u[2];
    const double s = fma(dz, dz, fma(dy, dy, dx * dx));
    const double d = sqrt(s);
    const double arg = (sign > 0 ? d : -d) * K;
    double c = cos(arg) / d, sn = sin(arg) / d;
    if (obliquity) {
        const double cf = fabs(dz) / d;
        c *= cf;
        sn *= cf;
    }
    *re = w_re * c - w_im * sn;
    *im = w_re * sn + w_im * c;
}

/* Brute-force sum (transform_full, transform.cu:147-340): for every target m
 * y[m] = sum_n pair(n, m), summed in source order.  u, v are AoS xyz
 * (macros.h:48-56), y is interleaved (re, im) like std::vector<WAVE>.
 * Targets listed in `targets` (length M_sel) are evaluated; pass NULL for
 * all M.  nthreads <= 0 uses the OpenMP default.  obliquity: see
 * oracle_pair. */
void oracle_transform_full(int sign, double lambda,
                           const double *amp, const double *phase,
                           const double *u, size_t N,
                           const double *v, size_t M,
                           const int64_t *targets, size_t M_sel,
                           double *y, int nthreads, int obliquity)
{
    const double K = oracle_wavenumber(lambda);
    const size_t count = targets ? M_sel : M;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (size_t j = 0; j < count; ++j) {
        const size_t m = targets ? (size_t)targets[j] : j;
        double sr = 0.0, si = 0.0;
        for (size_t n = 0; n < N; ++n) {
            double re, im;
            oracle_pair(sign, K, amp[n], phase[n], &u[3 * n], &v[3 * m],
                        obliquity, &re, &im, NULL, NULL);
            sr += re;
            si += im;
        }
        y[2 * j] = sr;
        y[2 * j + 1] = si;
    }
    (void)M;
}

/* Complex-weight brute force (adjoint direction): y[m] = sum_n w_n
 * exp(i sign d K)/d.  No reference counterpart; pinned in tests by a dense
 * complex-matrix product and finite differences. */
void oracle_transform_cartesian(int sign, double lambda,
                                const double *w /* interleaved, N */,
                                const double *u, size_t N,
                                const double *v, size_t M,
                                double *y, int nthreads, int obliquity)
{
    const double K = oracle_wavenumber(lambda);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (size_t m = 0; m < M; ++m) {
        double sr = 0.0, si = 0.0;
        for (size_t n = 0; n < N; ++n) {
            double re, im;
            oracle_pair_cartesian(sign, K, w[2 * n], w[2 * n + 1],
                                  &u[3 * n], &v[3 * m], obliquity, &re, &im);
            sr += re;
            si += im;
        }
        y[2 * m] = sr;
        y[2 * m + 1] = si;
    }
}

/* normalize_amp + polar conversion (transform.cu:106-139, 376, 400-405):
 * max_amp = max_m |y_m| over all entries; if max_amp < 1e-6 nothing is
 * divided; amp_m = |y_m| / max_amp, phase_m = atan2(im, re).
 * Returns max_amp.  Pinned by closed-form tests (max becomes 1, phase in
 * (-pi, pi]). */
double oracle_normalize_polar(const double *y, size_t M,
                              double *amp, double *phase)
{
    double max_amp = 0.0;
    for (size_t m = 0; m < M; ++m)
        max_amp = fmax(max_amp, hypot(y[2 * m], y[2 * m + 1]));
    const double div = max_amp < 1e-6 ? 1.0 : max_amp;
    for (size_t m = 0; m < M; ++m) {
        const double re = y[2 * m] / div, im = y[2 * m + 1] / div;
        amp[m] = hypot(re, im);
        phase[m] = atan2(im, re);
    }
    return max_amp;
}

/* ---------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Pinned by the Random123 known-answer vectors
 * and a pure-Python restatement in tests. */
static inline void philox_round(uint32_t c[4], const uint32_t k[2])
{
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                          uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    memcpy(out, c, sizeof c);
}

#define ORACLE_MC_STREAM 0x4d43u /* "MC": third counter word of index draws */

/* Sample index n_{m,k} in [0, N). */
uint64_t oracle_mc_index(uint64_t seed, uint64_t m, uint64_t k, uint64_t N)
{
    const uint32_t ctr[4] = {(uint32_t)(k >> 1), (uint32_t)m,
                             (uint32_t)(m >> 32), ORACLE_MC_STREAM};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    oracle_philox4x32_10(ctr, key, o);
    const uint64_t r = (k & 1) ? ((uint64_t)o[3] << 32 | o[2])
                               : ((uint64_t)o[1] << 32 | o[0]);
    return (uint64_t)(((unsigned __int128)r * N) >> 64);
}

/* Shared-subset ("Batched MC") index: group g's k-th draw, counter
 * (k/2, g, 0x80000000, 0x4d43). */
#define ORACLE_MC_SHARED_TAG 0x80000000u
#define ORACLE_MC_GROUP 1024u

uint64_t oracle_mc_shared_index(uint64_t seed, uint64_t g, uint64_t k, uint64_t N)
{
    const uint32_t ctr[4] = {(uint32_t)(k >> 1), (uint32_t)g, ORACLE_MC_SHARED_TAG,
                             ORACLE_MC_STREAM};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    oracle_philox4x32_10(ctr, key, o);
    const uint64_t r = (k & 1) ? ((uint64_t)o[3] << 32 | o[2])
                               : ((uint64_t)o[1] << 32 | o[0]);
    return (uint64_t)(((unsigned __int128)r * N) >> 64);
}

/* MC estimator in FP64 with the pinned pair: (N/K) sum_k pair(n_{m,k}, m).
 * shared != 0 uses the group subsets n_{m/1024, k}. */
void oracle_transform_mc(int sign, double lambda,
                         const double *amp, const double *phase,
                         const double *u, size_t N,
                         const double *v, size_t M,
                         uint32_t K_samples, uint64_t seed, int shared,
                         const int64_t *targets, size_t M_sel,
                         double *y, int nthreads, int obliquity)
{
    const double K = oracle_wavenumber(lambda);
    const size_t count = targets ? M_sel : M;
    const double scale = (double)N / (double)K_samples;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (size_t j = 0; j < count; ++j) {
        const size_t m = targets ? (size_t)targets[j] : j;
        double sr = 0.0, si = 0.0;
        for (uint32_t k = 0; k < K_samples; ++k) {
            const uint64_t n = shared ? oracle_mc_shared_index(seed, m / ORACLE_MC_GROUP, k, N)
                                      : oracle_mc_index(seed, m, k, N);
            double re, im;
            oracle_pair(sign, K, amp[n], phase[n], &u[3 * n], &v[3 * m],
                        obliquity, &re, &im, NULL, NULL);
            sr += re;
            si += im;
        }
        y[2 * j] = sr * scale;
        y[2 * j + 1] = si * scale;
    }
    (void)M;
}

/* Gradient of a real loss L(z) through z = sum_n A_n e^{i theta_n} G_mn,
 * G_mn = exp(i sign d K)/d, given g_m = dL/dRe z_m + i dL/dIm z_m:
 *   b_n        = sum_m g_m conj(G_mn)
 *   dL/dA_n    = Re(e^{-i theta_n} b_n)
 *   dL/dtheta_n = A_n Im(e^{-i theta_n} b_n)
 * No reference counterpart.  Pinned by central finite differences and a
 * dense complex-matrix autograd product in tests.  With obliquity, G carries
 * the factor |dz| / d, which is symmetric in (m, n), so b keeps the same
 * form. */
void oracle_adjoint(int sign, double lambda,
                    const double *amp, const double *theta,
                    const double *u, size_t N,
                    const double *g /* interleaved, M */,
                    const double *v, size_t M,
                    double *grad_amp, double *grad_theta, int nthreads,
                    int obliquity)
{
    double *b = (double *)malloc(sizeof(double) * 2 * N);
    /* b_n: sources are the targets v carrying weights g, targets are u. */
    oracle_transform_cartesian(-sign, lambda, g, v, M, u, N, b, nthreads, obliquity);
    for (size_t n = 0; n < N; ++n) {
        const double c = cos(theta[n]), s = sin(theta[n]);
        /* e^{-i theta} b = (c br + s bi) + i (c bi - s br) */
        const double pr = c * b[2 * n] + s * b[2 * n + 1];
        const double pi = c * b[2 * n + 1] - s * b[2 * n];
        grad_amp[n] = pr;
        grad_theta[n] = amp[n] * pi;
    }
    free(b);
}
