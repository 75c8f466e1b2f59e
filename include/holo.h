/*
 * holo.h -- C ABI of the MI355X (gfx950) Huygens superposition library
 * (libholo.so).
 *
 * The library computes the point-source superposition of
 * voschezang/Holographic-Projector-Simulations:
 *
 *     y_m = sum_n  a_n / d_nm * exp(i (phi_n + sign * 2 pi d_nm / lambda)),
 *     d_nm = |v_m - u_n|_2
 *
 * (cuda/superposition.cu:35-51), with sign = +1 for Direction::Forwards and
 * -1 for Direction::Backwards (cuda/superposition.cu:19).
 *
 * Conventions
 *  - Positions are AoS xyz doubles, element i at [3i, 3i+1, 3i+2]
 *    (cuda/macros.h:48-56 Ix(i, dim)).
 *  - Fields are interleaved complex doubles (re, im), like
 *    std::vector<cuDoubleComplex> (cuda/macros.h:40 WAVE).
 *  - Without HOLO_DEVICE_POINTERS every array is host memory and the call
 *    returns after the result is in host memory.  With it every array is
 *    device memory on the context's device and the call is ordered on the
 *    context's stream without synchronising.
 *  - Errors are returned as holo_status; the thread's last message is
 *    available from holo_last_error().  Nothing calls exit() or abort().
 *    Flag bits that a call does not accept are rejected with HOLO_EINVAL.
 *  - N and M must be below 2^31.  A source and a target at the same point
 *    (d = 0) yield a non-finite term, as in the reference.
 *  - One context per host thread; calls on one context are not reentrant.
 *  - Error report (err_rel_l2, nullable, of every call that returns a field
 *    or gradients): for HOLO_FP32_SPLIT the call also evaluates the FP64
 *    formulation r64 at count = min(M, 256) targets of its result r (M
 *    values) and writes an estimate of the relative L2 error of the whole
 *    result, sqrt(sum_m |r - r64|^2 / T) with T = sum_m |r|^2.  n_e =
 *    count / 2 targets are drawn by energy (draw i is the first target at
 *    which the running sum of |r|^2 exceeds (i + 1/2) / n_e of T), n_u =
 *    count - n_e are m_j = floor(j M / n_u), and the estimate is
 *    sqrt(sum_i |r[m_i] - r64_i|^2 / (n_e |r[m_i]|^2 + n_u T / M)); with
 *    M <= 256 every target is taken once (n_e = 0) and the value is exact.
 *    A result without a finite positive T reports the relative L2 error at
 *    the targets floor(i M / count).  For HOLO_FP64 it writes 0 (the result
 *    is the FP64 formulation).  The extra work is 256 N FP64 pairs (256 K
 *    samples for Monte Carlo) and two passes over r.  err_rel_l2 is a host
 *    pointer, or a device pointer with HOLO_DEVICE_POINTERS (the call then
 *    stays asynchronous).
 */
#ifndef HOLO_H
#define HOLO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HOLO_ABI_VERSION 2

typedef enum holo_status {
  HOLO_OK = 0,
  HOLO_EINVAL = 1,  /* invalid argument: size, sign, lambda, null pointer, precision */
  HOLO_ENOMEM = 2,  /* device or host allocation failed */
  HOLO_EDEVICE = 3, /* HIP runtime error or no usable gfx950 device */
  HOLO_EIO = 4      /* file read/write error (holo_read_dat / holo_write_dat) */
} holo_status;

typedef enum holo_precision {
  HOLO_FP64 = 0,      /* plain FP64: the drop-in; matches the FP64 CPU formulation */
  HOLO_FP32_SPLIT = 1 /* FP32 pair math on FP64 per-tile references (~3e-7 rel. L2) */
} holo_precision;

#define HOLO_FORWARDS (+1)  /* phase + 2 pi d / lambda (Direction::Forwards) */
#define HOLO_BACKWARDS (-1) /* phase - 2 pi d / lambda (Direction::Backwards) */

#define HOLO_DEVICE_POINTERS 0x1u /* all array arguments are device pointers */
#define HOLO_MC_SHARED 0x2u       /* Monte Carlo: groups of 1024 targets share a subset */
/* Obliquity-weighted pairs: every pair's amplitude is also multiplied by
 * cos = |v_z - u_z| / d, the `cosfact = Gz ./ Gdistances` of the MATLAB
 * light-field model (matlab/holofunctions.m:226-227, which assumes one plane
 * separation Gz for all pairs).  Accepted by holo_transform_full,
 * holo_transform_mc, holo_plan_transform(_complex) and holo_plan_adjoint. */
#define HOLO_OBLIQUITY 0x4u

typedef struct holo_ctx holo_ctx;
typedef struct holo_plan holo_plan;

/* ---- library / context ------------------------------------------------- */

int holo_abi_version(void);
const char *holo_status_string(holo_status status);
/* Message of the last failing call on this thread ("" if none). */
const char *holo_last_error(void);

/* Create a context on HIP device `device` using `stream` (a hipStream_t, or
 * NULL for a stream owned by the context).  Replaces the per-call stream and
 * cuBLAS-handle creation of cuda/transform.cu:233-241. */
holo_status holo_create(holo_ctx **ctx, int device, void *stream);
void holo_destroy(holo_ctx *ctx);
/* Use `stream` from now on; NULL selects the device's default (null) stream,
 * as torch's default stream is.  A stream the context created is released
 * (after a host wait for it).  Otherwise work later queued on `stream` waits
 * on the device for the work queued so far on the previous stream, which may
 * still use the context's scratch buffers; the host does not wait.  A stream
 * must stay valid while it is the context's stream. */
holo_status holo_set_stream(holo_ctx *ctx, void *stream);
holo_status holo_synchronize(holo_ctx *ctx);

/* Tuning knobs (process-wide): "r_f64", "r_f32" targets per lane (1, 2, 4),
 * "splits" source splits (0 = automatic), "target_waves", "f64_libm" (0/1:
 * use the device libm sqrt/div/sincos in the FP64 kernel), "f64_planar"
 * (0/1: when every source of a plan, or of a host-pointer transform, has the
 * same z, the FP64 kernel forms dz once per target; same bits; default 1),
 * "f64_rows" (0/1: with "f64_planar", when the sources also come in rows of
 * consecutive sources with one x, as lattices in x-slow order do, dx^2 is
 * formed once per target and row; same bits; default 1),
 * "grid" (0/1: HOLO_FP32_SPLIT on lattice sources uses the matrix-core
 * lattice kernel; default 1), "grid16" (0/1: on a uniform lattice that
 * kernel uses 16 x 16 tiles, padded to multiples of 16; default 1), "grid32"
 * (0/1: with "grid16", where the padded count of 16-column tile columns is
 * even, tiles of 16 x 32; default 1), "grid64" (0/1: with "grid32", where
 * that count is a multiple of 4, tiles of 16 x 64; default 1), "grid_series" (0/1: pairs of the 16-row
 * kernels too near for the GEMM factorisation are summed by the q-series to
 * third order where its bounds hold, else exactly; default 1), "grid_near"
 * (0/1: waves of targets off the 16-row kernels' far-field bounds take the
 * near-field GEMM, its cross factor to second order, before the q-series and
 * exact paths; default 1), "grid_rows" (0/1: the 16-row kernels' exact
 * pairs take the x of their row and the y of their column from tables of
 * the lattice's rows and columns, dx^2 once per row, instead of each
 * source's position; same bits; default 1), "mc_passes"
 * (per-target Monte Carlo: source ranges, one launch each, 0 = automatic,
 * 1 = one launch), "mc_lattice" (0/1: per-target
 * FP32-split Monte Carlo on product-grid sources reads positions from row
 * and column tables; same results; default 1).  Results do not depend on them beyond summation order and, for "grid",
 * "grid16", "grid32" and "grid64", the FP32-split rounding (all paths stay within
 * ~1e-6).  Replaces the blockDim/gridDim/thread_size/n_streams hints of
 * cuda/input.h:116-122. */
holo_status holo_set_option(const char *name, long value);
long holo_get_option(const char *name);

/* ---- brute force: transform_full ----------------------------------------
 * Replaces transform_full<Direction, Algorithm, shared_memory>(x, u, v,
 * Geometry) (cuda/transform.cu:147-340): y[m] for m < M, unnormalised.
 * amp, phase: N doubles; u: 3N; v: 3M; y: 2M (interleaved complex);
 * err_rel_l2: the error report above (may be NULL). */
holo_status holo_transform_full(holo_ctx *ctx, holo_precision precision,
                                int sign, double lambda, const double *amp,
                                const double *phase, const double *u,
                                size_t N, const double *v, size_t M,
                                double *y, double *err_rel_l2, unsigned flags);

/* ---- no-allocation variant of holo_transform_full -------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops and harnesses that
 * forbid hipMalloc.  Computes what holo_transform_full computes with
 * HOLO_FP64 and device pointers (all arrays are device memory on the current
 * device; HOLO_DEVICE_POINTERS may be given and changes nothing;
 * HOLO_OBLIQUITY is accepted), ordered on `stream` (a hipStream_t of the
 * current device, NULL = its null stream) without synchronising.  It needs no
 * context and calls no allocating routine: its scratch (the source-split
 * partial fields and four words of source layout) is a static device array of
 * HOLO_NOALLOC_SCRATCH_BYTES per device, held from the library's first use
 * on a device (holo_create, or the first call of this function there) until
 * the process ends.  The call uses at most (HOLO_NOALLOC_SCRATCH_BYTES - 256)
 * / (16 M) source splits, so every M below 2^31 fits (one split writes y
 * directly).  Every call on a device shares that device's array: calls that
 * overlap on one device (other streams, other threads) race and must be
 * ordered by the caller.  Results equal holo_transform_full's to FP64
 * rounding of the split sums (bit for bit when both use the same splits). */
#define HOLO_NOALLOC_SCRATCH_BYTES (32u << 20)
holo_status holo_transform_full_noalloc(int sign, double lambda, const double *amp,
                                        const double *phase, const double *u, size_t N,
                                        const double *v, size_t M, double *y,
                                        void *stream, unsigned flags);

/* ---- Monte Carlo: transform ---------------------------------------------
 * The reference's transform() (named in README.md:127, absent from the
 * snapshot; remnants cuda/kernel.cu:101-106,197-212): each target m draws K
 * source indices n_{m,k} uniformly with replacement and
 *     y[m] = (N / K) sum_k pair(n_{m,k}, m).
 * Indices come from Philox4x32-10 keyed by `seed`, counter
 * (k/2, m_lo, m_hi, 0x4d43); words (0,1) serve even k, (2,3) odd k; the
 * 64-bit word r maps to floor(r N / 2^64).  They depend only on
 * (seed, m, k, N), not on the launch.
 * With HOLO_MC_SHARED ("Batched MC", py/plot.py:787-804): the targets of
 * group g = m / 1024 share the subset n_{g,k} drawn with counter
 * (k/2, g, 0x80000000, 0x4d43); each target's estimate is still unbiased
 * with the same variance, but errors are correlated within a group.
 * err_rel_l2 (may be NULL): the error report above against the FP64
 * estimator with the same indices (not the statistical error, which is
 * that of the estimator itself). */
holo_status holo_transform_mc(holo_ctx *ctx, holo_precision precision,
                              int sign, double lambda, const double *amp,
                              const double *phase, const double *u, size_t N,
                              const double *v, size_t M, uint32_t K,
                              uint64_t seed, double *y, double *err_rel_l2,
                              unsigned flags);

/* Sample indices n_{m,k} for targets m0 .. m0+count-1 (row-major count x K,
 * int64), as drawn on the device by holo_transform_mc; with HOLO_MC_SHARED,
 * the subsets of groups m0 .. m0+count-1. */
holo_status holo_mc_indices(holo_ctx *ctx, uint64_t seed, uint64_t N,
                            size_t m0, size_t count, uint32_t K,
                            int64_t *idx, unsigned flags);

/* ---- time_transform post-processing --------------------------------------
 * normalize_amp + polar conversion of cuda/transform.cu:106-139, 376,
 * 400-405: max_amp = max |y|; if max_amp >= 1e-6, y /= max_amp; then
 * amp = |y|, phase = atan2(im, re).  *max_amp receives the maximum (host
 * pointer, or device pointer with HOLO_DEVICE_POINTERS; may be NULL). */
holo_status holo_normalize_polar(holo_ctx *ctx, const double *y, size_t M,
                                 double *amp, double *phase, double *max_amp,
                                 unsigned flags);

/* ---- plans: a fixed source geometry reused across calls -----------------
 * A plan keeps the N source positions resident on the device and, for
 * HOLO_FP32_SPLIT, groups them into spatial tiles of `tile` sources
 * (8, 16, 32 or 64; 0 = 64).  No reference counterpart: it replaces the
 * per-call host-to-device copies of cuda/transform.cu:209-213.
 * Each call rewrites the plan's weight buffers.  A plan may serve several
 * contexts: a call on a stream other than that of the plan's previous call
 * waits on the device for that call.  Calls must not overlap on the host. */
holo_status holo_plan_create(holo_ctx *ctx, holo_plan **plan, const double *u,
                             size_t N, unsigned tile, unsigned flags);
void holo_plan_destroy(holo_plan *plan);
/* n_sources, n_tiles, tile, and the largest source offset from its tile
 * centre (metres); any pointer may be NULL. */
holo_status holo_plan_info(const holo_plan *plan, size_t *n_sources,
                           size_t *n_tiles, unsigned *tile,
                           double *max_offset);

/* Lattice the plan detected: the sources are a product grid x_i * y_j at one
 * z, in the order n = i*ny + j (the reference's plane generators,
 * cuda/init.h:89-158), either with nx and ny multiples of 8, or uniform
 * (x_0 + i p_x, y_0 + j p_y) with nx, ny >= 16 and followed by fewer than
 * max(nx, ny) other sources (the generator fills a point count that is not
 * nx * ny with copies of source 0, cuda/init.h:153-157).  HOLO_FP32_SPLIT then
 * runs the lattice kernel: tiles of 16 x 32 or 16 x 16 sources on a uniform
 * lattice (padded with weight-0 slots to multiples of 16, the sources after
 * it summed exactly by every target), 8 x 8 on other product grids (and with
 * option "grid16" = 0), each summed as a complex GEMM on the matrix cores.  nx = ny = 0 when the plan is not a lattice.
 * No reference counterpart. */
holo_status holo_plan_lattice(const holo_plan *plan, size_t *nx, size_t *ny);

/* Host-only preview of the tiling a plan builds (no device needed): tile
 * count, slot -> source index map (n_tiles * tile entries, -1 = padding;
 * perm may be NULL), largest offset from a tile centre, and whether every
 * tile lies in one z plane. */
holo_status holo_plan_layout(const double *u, size_t N, unsigned tile,
                             int32_t *perm, size_t perm_capacity,
                             size_t *n_tiles, double *max_offset, int *planar);

/* Host-only preview of the lattice a plan with the default tile detects (no
 * device needed): nx, ny as holo_plan_lattice reports them (0, 0: not a
 * lattice), the number of sources after the lattice that every target sums
 * exactly (tail), and whether the 16-row tiles are built (tiles16).  Any
 * output pointer may be NULL.  No reference counterpart. */
holo_status holo_lattice_layout(const double *u, size_t N, size_t *nx, size_t *ny,
                                size_t *tail, int *tiles16);

/* Polar weights (amp, phase: N each), as holo_transform_full. */
holo_status holo_plan_transform(holo_ctx *ctx, const holo_plan *plan,
                                holo_precision precision, int sign,
                                double lambda, const double *amp,
                                const double *phase, const double *v, size_t M,
                                double *y, double *err_rel_l2, unsigned flags);

/* Complex weights w (interleaved, N): y[m] = sum_n w_n exp(i sign d K)/d. */
holo_status holo_plan_transform_complex(holo_ctx *ctx, const holo_plan *plan,
                                        holo_precision precision, int sign,
                                        double lambda, const double *w,
                                        const double *v, size_t M, double *y,
                                        double *err_rel_l2, unsigned flags);

/* Gradient of a real loss L through z = transform(amp, theta) with sources
 * u (N) and targets the plan's points (M): given
 * g_m = dL/dRe z_m + i dL/dIm z_m (interleaved, M),
 *     b_n            = sum_m g_m conj(G_mn),  G_mn = exp(i sign d K)/d
 *     grad_amp[n]    = Re(exp(-i theta_n) b_n)
 *     grad_theta[n]  = amp_n Im(exp(-i theta_n) b_n)
 * `plan_targets` tiles the forward targets (they are the adjoint's sources);
 * `sign` is the forward sign.  err_rel_l2 (may be NULL): the error report
 * above of the field b at the N points u.  No reference counterpart. */
holo_status holo_plan_adjoint(holo_ctx *ctx, const holo_plan *plan_targets,
                              holo_precision precision, int sign,
                              double lambda, const double *g,
                              const double *amp, const double *theta,
                              const double *u, size_t N, double *grad_amp,
                              double *grad_theta, double *err_rel_l2,
                              unsigned flags);

/* ---- .dat files -----------------------------------------------------------
 * Headerless little-endian float64 arrays (cuda/util.h:248-283 read_bytes,
 * 295-337 write_bytes).  holo_read_dat: *count receives the number of
 * doubles; with buf == NULL only the count is returned. */
holo_status holo_read_dat(const char *path, double *buf, size_t capacity,
                          size_t *count);
holo_status holo_write_dat(const char *path, const double *buf, size_t count);

/* ---- host builds of the device math (verification only) -----------------
 * The FP64 kernels' sin/cos (Cody-Waite reduction modulo pi/1024, a table of
 * (sin, cos)(j pi/1024) and short polynomials)
 * and sqrt / reciprocal, compiled for the host from the same source; the
 * sin/cos bits match the device's.  No reference counterpart. */
void holo_host_sincos(double x, double *s, double *c);
void holo_host_sqrt_inv(double s, double *d, double *inv);

#ifdef __cplusplus
}
#endif

#endif /* HOLO_H */
