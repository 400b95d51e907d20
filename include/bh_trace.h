/*
 * bh_trace.h -- C ABI of the MI355X (gfx950) batched adaptive Kerr geodesic
 * tracer.  All compute runs in hand-written HIP kernels, FP64 throughout.
 *
 * Conventions
 *  - Every buffer is device memory owned by the caller (for example a
 *    PyTorch-ROCm tensor passed by data_ptr()).  The library never allocates
 *    caller-visible memory and never frees caller memory.
 *  - Work is enqueued on the caller's HIP stream (`stream`, a hipStream_t;
 *    NULL means the default stream).  No call synchronises the device.
 *  - Return value: BH_OK (0) or a negative BH_E* code.  bh_last_error()
 *    returns a thread-local message describing the last failure.
 *  - Ray counts (n, width * height) are at most 2^36 per call; larger
 *    counts, negative counts and ld < n are BH_EINVAL.
 *  - "state" layout at the boundary: rows (t, x, y, z, p_t, p_x, p_y, p_z)
 *    with covariant momenta (reference geodesics.py:3-8).  The tracer
 *    kernels take structure-of-arrays planes: component c of ray i lives at
 *    base[c * ld + i], ld >= n.  bh_aos_to_soa_f64/bh_soa_to_aos_f64 convert
 *    from/to the (n, 8) row-major layout of the Python API.
 *  - Ray status codes equal the reference's RayStatus (tracer.py:21-30).
 *
 * Citations are file:line in the reference repository (MegaKruk/BlackHorizon).
 */
#ifndef BH_TRACE_H
#define BH_TRACE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_ABI_VERSION 4  /* 2: integrate_dense work is 24 ld + 8192 doubles; static-scratch adjoint;
                             3: bh_tune_params.lanes_per_ray (quad tracer);
                             4: trajectory-store record / reverse pass, radial minimum */

#define BH_OK 0
#define BH_EINVAL -1   /* invalid argument (maps to ValueError) */
#define BH_EHIP -2     /* HIP runtime / launch failure (maps to RuntimeError) */

/* tracer.py:21-30 */
#define BH_IN_FLIGHT 0
#define BH_CAPTURED 1
#define BH_ESCAPED 2
#define BH_MAX_STEPS 3
#define BH_FAILED 4
#define BH_DISK 5
#define BH_TERMINATED 6

/* Arithmetic mode of a kernel.
 * BH_MODE_FAST:    FMA contraction, shared reciprocals, squared error ratio
 *                  with a polynomial-corrected step factor.
 * BH_MODE_PRECISE: one correctly rounded IEEE operation per NumPy ufunc of
 *                  the reference, in the reference's order (no contraction),
 *                  x ** -0.2 correctly rounded; used to show rounding-level
 *                  parity. */
#define BH_MODE_FAST 0
#define BH_MODE_PRECISE 1

/* Controller settings of one trace job: the keyword arguments of
 * trace_rays (tracer.py:50-61) plus the spacetime (kerr.py:61-73).
 * capture_radius is r_+ (1 + capture_margin), computed by the caller exactly
 * as tracer.py:84 does, so that the comparison at tracer.py:128-132 uses the
 * same double. */
typedef struct bh_trace_params {
  double mass;
  double spin;
  double escape_radius;
  double capture_radius;
  double rtol;
  double atol;
  double initial_step;
  double max_step;
  double min_step;
  int64_t max_steps;   /* accepted-step cap; attempts are capped at 4 * max_steps */
  int32_t mode;        /* BH_MODE_FAST or BH_MODE_PRECISE */
  int32_t reserved;
  /* Affine-parameter end; <= 0, NaN or +inf: none (so a zero-initialised
   * struct traces like trace_rays).  Each attempt uses
   * min(h, lambda_end - lambda) as integrate_dense does
   * (integrators.py:161-162); a ray reaching lambda_end ends with status
   * BH_TERMINATED.  For timelike rays with H = -1/2 lambda is proper time. */
  double lambda_end;
} bh_trace_params;

/* Launch shape of the tracers.  Zero fields pick defaults. */
typedef struct bh_tune_params {
  int32_t block_size;      /* threads per block, multiple of 64 (default 256) */
  int32_t rays_per_lane;   /* plain tracer: rays each lane traces in turn (>= 1; default 1) */
  int32_t refill_min;      /* idle lanes (quads with lanes_per_ray 4) a wave collects
                              before refilling (1..64; 1..16 for quads) */
  int32_t lanes_per_ray;   /* batched tracer: 1 (default) or 4 -- the four lanes of a
                              DPP quad trace one ray (precise build only), for launches
                              whose time is one long trace's latency; the plain tracer
                              takes 1 only.  Results are the same either way. */
} bh_tune_params;

int32_t bh_abi_version(void);
const char *bh_last_error(void);

/* Resource usage of one tracing kernel of one arithmetic build, for reports:
 * out[0] registers per thread (VGPRs), out[1] static LDS bytes per block,
 * out[2] private scratch bytes per thread (spills), out[3] resident blocks
 * per CU at `block` threads per block.  No reference counterpart. */
#define BH_KERNEL_PLAIN 0
#define BH_KERNEL_BATCHED 1
#define BH_KERNEL_RECORD 2
#define BH_KERNEL_ADJOINT 3
#define BH_KERNEL_TILE 4
#define BH_KERNEL_SHADER 5
#define BH_KERNEL_QUAD 6    /* precise build only */
int32_t bh_kernel_attributes(int32_t kernel, int32_t mode, int32_t block, int32_t *out);

/* Number of rays with r0 >= escape_radius, written to *count_out (device
 * int32).  Replaces the initial-radius check of tracer.py:94-99; the host
 * raises ValueError when the count is nonzero. */
int32_t bh_count_outside_f64(const bh_trace_params *p, const double *state_soa,
                             int64_t ld, int64_t n, int32_t *count_out,
                             void *stream);

/* Trace n rays to termination: one DP5(4) attempt per loop pass per ray,
 * exactly the controller of tracer.py:101-154 applied per ray.  Outputs:
 * final states (SoA), status (int32), accepted steps (int64), attempts
 * (int64, accepted + rejected), and *iterations (device uint64) = max
 * attempts over the rays (tracer.py:155-157 `iterations`).
 * order: optional device int64[n] permutation of [0, n): rays are handed to
 * wavefront lanes in this order (results still land at each ray's own
 * index); NULL = identity.  Results do not depend on order or tune.
 * Replaces trace_rays (tracer.py:50-157). */
int32_t bh_trace_plain_f64(const bh_trace_params *p, const bh_tune_params *tune,
                           const double *state_in_soa, int64_t ld_in,
                           double *state_out_soa, int64_t ld_out,
                           int32_t *status, int64_t *steps, int64_t *attempts,
                           uint64_t *iterations, const int64_t *order, int64_t n,
                           void *stream);

/* Persistent batched trace of n_jobs independent jobs in one launch.
 * jobs: device array of n_jobs bh_trace_params (each job's own spacetime,
 *   camera-independent controller settings; the `mode` field is ignored --
 *   the launch's `mode` selects the arithmetic build for all jobs).  Jobs must
 *   satisfy the checks bh_trace_plain_f64 applies to its params.
 * ray_offsets: device int64[n_jobs + 1], nondecreasing, ray_offsets[0] = 0,
 *   ray_offsets[n_jobs] = n_total: job j owns rays [off[j], off[j + 1]) of
 *   the concatenated SoA planes.
 * iterations: device uint64[n_jobs], per-job max attempts (tracer.py:155-157).
 * order: optional device int64[n_total] permutation: the queue hands out
 *   ray order[k] k-th (NULL = job order).
 * workspace: device uint64[1] scratch (the ray-queue counter).
 * Outputs per ray exactly as bh_trace_plain_f64 would give for each job
 * alone; only the schedule differs (lanes refill from a global ray queue as
 * rays retire).  n_jobs <= 256.  tune->lanes_per_ray 4 (precise mode only)
 * traces each ray on the four lanes of a DPP quad, one wavefront per SIMD:
 * same outputs, shorter time per attempt for launches that end on a few long
 * rays.  Replaces a loop of trace_rays calls (tracer.py:50-157), one per
 * job. */
int32_t bh_trace_batched_f64(const bh_trace_params *jobs, const int64_t *ray_offsets,
                             int32_t n_jobs, int32_t mode, const bh_tune_params *tune,
                             const double *state_in_soa, int64_t ld_in,
                             double *state_out_soa, int64_t ld_out, int32_t *status,
                             int64_t *steps, int64_t *attempts, uint64_t *iterations,
                             const int64_t *order, uint64_t *workspace, int64_t n_total,
                             void *stream);

/* ---- discrete adjoint (differentiable tracing) -------------------------
 * No counterpart in the reference (it has no derivatives); these extend
 * trace_rays (tracer.py:50-157) to gradients of each ray's terminal point
 * with respect to its initial state and the spin. */

/* Doubles of adjoint scratch per ray for checkpoint interval K (9 K),
 * or -1 for K < 1. */
int32_t bh_adjoint_scratch_doubles(int32_t K);

/* Forward trace exactly as bh_trace_plain_f64 (same statuses, steps and
 * final states; one ray per lane), recording for the adjoint:
 *  ckpt  [cap][10][n]: state, controller step size and affine parameter
 *  every K accepted steps;
 *  n_ckpt [n]; term [8][n]: terminal point -- escaped rays: where the cubic
 *  Hermite dense output of the last step crosses the escape sphere;
 *  captured rays: where the step that first enters the sphere
 *  r = score_radius crosses it; otherwise the final state;
 *  prev [9][n] / evy1 [8][n]: start state + step size / end state of that
 *  event step; ev_step [n]: its accepted-step index; theta [n]: its dense-
 *  output parameter; event [n]: 0 none, 1 escape, 2 scoring sphere,
 *  -1 checkpoint capacity exceeded (no gradient).
 * order: optional device int64[n] permutation, lane slot k traces ray
 * order[k] (NULL = identity); rays of similar length in one wavefront
 * finish together.  Results land at each ray's own index.
 * workspace: device uint64[1] scratch (a ray-queue counter): a precise-mode
 * recording of at most one wavefront per SIMD runs on DPP quads of lanes,
 * one ray per quad, handed out in `order` -- the same outputs, bit for bit,
 * sooner when a few long rays set the time. */
int32_t bh_trace_record_f64(const bh_trace_params *p, const double *state_in_soa, int64_t ld_in,
                            double *state_out_soa, int64_t ld_out, int32_t *status,
                            int64_t *steps, int64_t *attempts, uint64_t *iterations,
                            const int64_t *order, int64_t n, int32_t K, int32_t cap,
                            double score_radius, double *ckpt,
                            int32_t *n_ckpt, double *term, double *prev, double *evy1,
                            int64_t *ev_step, double *theta, int32_t *event, uint64_t *workspace,
                            void *stream);

/* Accepted steps per page of a trajectory store, and the longest run of
 * consecutive pages a ray takes from the pool at once (its runs are 1, 2,
 * 4, 8, ... pages, then BH_TRAJ_RUN_MAX each). */
#define BH_TRAJ_PAGE_STEPS 32
#define BH_TRAJ_RUN_MAX 16

/* bh_trace_record_f64 storing every accepted step instead of checkpoints,
 * for bh_trace_adjoint_trajectory_f64 (no replay in the reverse pass).  The
 * store is a pool of pool_pages pages of BH_TRAJ_PAGE_STEPS steps (pool:
 * pool_pages * BH_TRAJ_PAGE_STEPS * 9 device doubles); a ray opens a page
 * when its accepted step s is a multiple of BH_TRAJ_PAGE_STEPS -- the next
 * page of the run of consecutive pages it last took from the pool through
 * an atomic counter, or, with that run used up, the first of a new run
 * (runs of 1, 2, 4, ... pages, at most BH_TRAJ_RUN_MAX) -- and writes step
 * s (start state, then the accepted step size) at pool[(page *
 * BH_TRAJ_PAGE_STEPS + s mod BH_TRAJ_PAGE_STEPS) * 9 + 0..8].  page_link [pool_pages]: each taken
 * page's predecessor of the same ray (-1 for its first page); last_page [n]:
 * each ray's last page (-1 without steps).  workspace: device uint64[2]
 * (the quad form's ray queue, the pool counter); after the call
 * workspace[1] holds the pages requested (runs included in full): above
 * pool_pages, the rays whose run did not fit stopped storing and have
 * event -1.  Every output
 * other than the store is bh_trace_record_f64's, bit for bit. */
int32_t bh_trace_record_trajectory_f64(
    const bh_trace_params *p, const double *state_in_soa, int64_t ld_in, double *state_out_soa,
    int64_t ld_out, int32_t *status, int64_t *steps, int64_t *attempts, uint64_t *iterations,
    const int64_t *order, int64_t n, double *pool, int64_t pool_pages, int64_t *page_link,
    int64_t *last_page, double score_radius, double *term, double *prev, double *evy1,
    int64_t *ev_step, double *theta, int32_t *event, uint64_t *workspace, void *stream);

/* Reverse pass: given gterm [8][n] = dL/d(terminal point), writes
 * gy0 [8][n] = dL/d(initial state) and gspin [n] = per-ray dL/d(spin),
 * differentiating every accepted step with its step size held fixed (each
 * checkpoint segment is replayed through the controller).  scratch:
 * bh_adjoint_scratch_doubles(K) * (n rounded up to a multiple of 64) device
 * doubles (opaque; laid out per wavefront of 64 rays).  order: optional
 * device int64[n] permutation, lane slot k reverses ray order[k] (NULL =
 * identity); ordering rays by ev_step keeps a wavefront's lanes busy for
 * the same number of steps.  Rays with event -1 get NaN. */
int32_t bh_trace_adjoint_f64(const bh_trace_params *p, int64_t n, int32_t K, double score_radius,
                             const double *ckpt, const int32_t *event, const int64_t *ev_step,
                             const double *evy1, const double *prev, const double *theta,
                             const double *gterm, double *gy0, double *gspin, double *scratch,
                             const int64_t *order, void *stream);

/* Doubles in the static scratch of bh_trace_adjoint_static_f64 (64 MiB per
 * GPU). */
#define BH_ADJOINT_STATIC_SCRATCH_DOUBLES (8 * 1024 * 1024)

/* bh_trace_adjoint_f64 without the scratch argument: allocates nothing (no
 * hipMalloc, hipFree or allocating library routine), for callers that
 * cannot allocate -- code under HIP graph capture, tight fitting or serving
 * loops, harnesses that forbid allocation.  No counterpart in the
 * reference.  The scratch is a file-scope device array of
 * BH_ADJOINT_STATIC_SCRATCH_DOUBLES doubles, one per GPU, held until the
 * process ends; its device address is resolved on the first call on each
 * device (make that first call outside graph capture).  Limit: 9 K doubles
 * per ray with n rounded up to a multiple of 64 must fit
 * (9 K ceil(n / 64) 64 <= 2^23, e.g. n <= 29,120 rays at K = 32); larger
 * calls return BH_EINVAL.  Calls that overlap on one GPU share the array
 * and race: serialise them (one stream, or events).  Results equal
 * bh_trace_adjoint_f64's bit for bit. */
int32_t bh_trace_adjoint_static_f64(const bh_trace_params *p, int64_t n, int32_t K,
                                    double score_radius, const double *ckpt,
                                    const int32_t *event, const int64_t *ev_step,
                                    const double *evy1, const double *prev,
                                    const double *theta, const double *gterm, double *gy0,
                                    double *gspin, const int64_t *order, void *stream);

/* bh_trace_adjoint_f64 over a trajectory store written by
 * bh_trace_record_trajectory_f64 (the same pool, pool_pages, page_link and
 * last_page, and its steps output): each step's start state and size are
 * read instead of replayed, so there is no scratch and no checkpoint
 * interval.  Results equal bh_trace_adjoint_f64's on a checkpoint recording
 * of the same rays, bit for bit.  Rays with event -1, or whose page chain
 * leaves [0, pool_pages) before their event step, get NaN. */
int32_t bh_trace_adjoint_trajectory_f64(const bh_trace_params *p, int64_t n,
                                        double score_radius, const double *pool,
                                        int64_t pool_pages, const int64_t *page_link,
                                        const int64_t *last_page, const int64_t *steps,
                                        const int32_t *event, const int64_t *ev_step,
                                        const double *evy1, const double *prev,
                                        const double *theta, const double *gterm, double *gy0,
                                        double *gspin, const int64_t *order, void *stream);

/* Vector-Jacobian product of geodesic_rhs (geodesics.py:42-74) at n states
 * (rows): zbar = J^T v (n, 8) and abar = v . d f / d spin (n). */
int32_t bh_rhs_vjp_f64(double mass, double spin, const double *state, const double *v,
                       int64_t n, double *zbar, double *abar, void *stream);

/* The detached search of the shadow-fit loss's capture margin
 * (adjoint.capture_margin; no counterpart in the reference): per ray, the
 * minimiser r_star[i] of the normalised photon radial potential
 * R(r) / (r^2 + a^2)^2 with constants lam[i] = L_z / E, eta[i] = Q / E^2,
 * over [r_+, max(r_hi[i], r_+)], r_+ = mass + sqrt(mass^2 - a^2) -- the
 * first minimum on a grid of `grid` points, then `refine` golden-section
 * steps on the bracket of its grid neighbours; at_lo[i] = 1 when the grid
 * minimum is the first point.  spin: device pointer to the spin a (read on
 * the device, so the caller need not wait for it; |a| > mass gives NaN). */
int32_t bh_radial_minimum_f64(double mass, const double *spin, const double *lam,
                              const double *eta, const double *r_hi, int64_t n, int32_t grid,
                              int32_t refine, double *r_star, int32_t *at_lo, void *stream);

/* (n, 8) row-major <-> 8 planes of ld.  Replaces nothing in the reference
 * (its arrays stay (n, 8)); the boundary transpose of the SoA design. */
int32_t bh_aos_to_soa_f64(const double *aos, double *soa, int64_t ld, int64_t n,
                          void *stream);
int32_t bh_soa_to_aos_f64(const double *soa, int64_t ld, double *aos, int64_t n,
                          void *stream);

/* KerrSpacetime.geometry(x, y, z, gradients) (kerr.py:118-193) at n points
 * given as (n, 3) rows.  Outputs (any may be NULL when not wanted):
 * radius (n), h (n), l (n, 3), grad_h (n, 3), grad_l (n, 3, 3) with
 * grad_l[i][j] = d l_j / d x_i. */
int32_t bh_geometry_f64(double mass, double spin, int32_t mode,
                        const double *xyz, int64_t n, double *radius, double *h,
                        double *l, double *grad_h, double *grad_l, void *stream);

/* geodesic_rhs (geodesics.py:42-74): (n, 8) rows -> (n, 8) rows. */
int32_t bh_geodesic_rhs_f64(double mass, double spin, int32_t mode,
                            const double *state, int64_t n, double *deriv,
                            void *stream);

/* error_ratio (integrators.py:110-117) of n rows of dim (1..128) values of
 * state, new_state and error, each (n, dim) row-major: ratio (n) =
 * sqrt(mean((error / (atol + rtol max(|state|, |new_state|)))^2)),
 * non-finite -> +inf.  Precise: NumPy's operations and summation order, bit
 * for bit. */
int32_t bh_error_ratio_f64(int32_t mode, const double *state, const double *new_state,
                           const double *error, int64_t n, int32_t dim, double rtol,
                           double atol, double *ratio, void *stream);

/* step_factor (integrators.py:120-124) of n ratios: clip(0.9 max(ratio,
 * 1e-16)^-0.2, 0.2, 5).  Precise: the power correctly rounded; fast: the
 * fast build's power of the squared ratio (float estimate, one second-order
 * correction). */
int32_t bh_step_factor_f64(int32_t mode, const double *ratio, int64_t n, double *factor,
                           void *stream);

/* dormand_prince_step (integrators.py:76-107) with the geodesic RHS, then
 * error_ratio (110-117) and step_factor (120-124): (n, 8) rows and h (n) ->
 * new_state (n, 8), error (n, 8), ratio (n), factor (n).  ratio/factor may
 * be NULL. */
int32_t bh_dp5_step_f64(double mass, double spin, int32_t mode, double rtol,
                        double atol, const double *state, const double *h,
                        int64_t n, double *new_state, double *error,
                        double *ratio, double *factor, void *stream);

/* ---- offline tile tracer ---------------------------------------------
 * The reference's offline renderer traces with its own loop
 * (offline/render.py:184-309): each pass classifies the current state
 * first (capture at r <= capture_radius, escape at r >= escape_radius,
 * MAX_STEPS at steps >= max_steps), then attempts a DP5(4) step
 * (atol = rtol * 1e-3 is the caller's, initial h = initial_step = 0.1 there;
 * no min_step test); an accepted step whose z changes sign is refined by 40
 * bisections along an RK4 sub-step (render.py:312-349) and ends the ray as
 * BH_DISK when the crossing radius lies in [disk_inner, disk_outer]. */
#define BH_TILE_DISK 1     /* detect equatorial disk crossings */
#define BH_TILE_INTERIOR 2 /* interior camera: capture_radius is the stop
                            * radius, captures end BH_TERMINATED, and rays
                            * with |p| > 1e3, or |p| > 50 within 1.002 x the
                            * stop radius, stop too (render.py:236-246) */
typedef struct bh_tile_params {
  double disk_inner;
  double disk_outer;
  int64_t saturate_steps; /* steps counted as saturated: int(0.85 * max_steps) */
  int32_t flags;          /* BH_TILE_* */
  int32_t reserved;
} bh_tile_params;

/* Trace n rays with the offline loop.  Outputs per ray: status, accepted
 * steps, attempts, saturated (uint8: MAX_STEPS or steps >= saturate_steps),
 * escape_dir (n, 3): unit flat-space direction of escaped rays
 * (render.py:164-181 with p_t from the state), 0 otherwise; hit_pos /
 * hit_mom (n, 3): crossing position and spatial momentum of disk hits, 0
 * otherwise; optional final states (SoA, NULL to skip); *iterations (device
 * uint64) = max attempts.  order as bh_trace_plain_f64.  Replaces
 * _trace_tile (render.py:184-309) and _refine_crossings (312-349). */
int32_t bh_trace_tile_f64(const bh_trace_params *p, const bh_tile_params *t,
                          const bh_tune_params *tune, const double *state_in_soa,
                          int64_t ld_in, double *state_out_soa, int64_t ld_out,
                          int32_t *status, int64_t *steps, int64_t *attempts,
                          uint8_t *saturated, double *escape_dir, double *hit_pos,
                          double *hit_mom, uint64_t *iterations, const int64_t *order,
                          int64_t n, void *stream);

/* ---- the real-time shader's algorithm ---------------------------------
 * trace_like_shader (realtime/reference.py:41-188): fixed RK4 steps,
 * h = clip(step_scale (r - r_plus), min_step, max_step) (interior cameras:
 * step_scale min(max(|r - r_plus|, 0.15 r), 0.5 r)), then
 * h = min(h, 1 / max(1, |p|)); each pass first classifies: capture at
 * r <= capture_radius or |p|^2 >= momentum_bailout^2 (interior: status
 * TERMINATED, also r <= 1.002 capture_radius with |p|^2 > 2500), escape at
 * r >= escape_radius; an equatorial crossing interpolated linearly inside
 * [disk_inner, disk_outer] ends the ray BH_DISK; rays left after max_steps
 * passes stay BH_MAX_STEPS. */
typedef struct bh_shader_params {
  double mass;
  double spin;
  double capture_radius;  /* r_plus (1 + capture_margin), or the interior stop */
  double escape_radius;
  double step_scale;
  double min_step;
  double max_step;
  double momentum_bailout;
  double disk_inner;
  double disk_outer;
  int64_t max_steps;      /* the caller applies the interior cap of 1024 */
  int32_t flags;          /* BH_TILE_DISK | BH_TILE_INTERIOR */
  int32_t mode;
} bh_shader_params;

/* Outputs: final states (SoA, NULL to skip), status, and optional hit
 * (n, 8) = (crossing radius, position (3), momentum (4)) of disk hits, NaN
 * otherwise.  order as bh_trace_plain_f64. */
int32_t bh_trace_shader_f64(const bh_shader_params *p, const double *state_in_soa, int64_t ld_in,
                            double *state_out_soa, int64_t ld_out, int32_t *status, double *hit,
                            const int64_t *order, int64_t n, void *stream);

/* rk4_step (integrators.py:57-73) with the geodesic RHS, applied n_steps
 * times per ray (the reference's fixed-step throughput benchmark,
 * examples/benchmark.py:34-62, loops rk4_step n_steps times): in place on
 * state_soa [8][ld], per-ray step h (n). */
int32_t bh_rk4_f64(double mass, double spin, int32_t mode, double *state_soa, int64_t ld,
                   const double *h, int64_t n, int64_t n_steps, void *stream);

/* implicit_midpoint_step (integrators.py:177-207) with the geodesic RHS,
 * `iterations` fixed-point sweeps, applied n_steps times per ray, in place
 * on state_soa [8][ld]; per-ray step h (n).  Symplectic: bounded energy
 * error over long orbits. */
int32_t bh_midpoint_f64(double mass, double spin, int32_t mode, double *state_soa, int64_t ld,
                        const double *h, int64_t n, int64_t n_steps, int32_t iterations,
                        void *stream);

/* integrate_dense (integrators.py:127-174) with the geodesic RHS, one
 * resumable chunk: all n rays take the same step, set by the largest error
 * ratio of the batch; h = min(h, max_step, lambda_total - elapsed) before
 * each attempt, h *= step_factor(max ratio) after it.  In/out state
 * state_soa [8][ld] (the states at the chunk's start, then at its end) and
 * ctl [2] = {elapsed affine parameter, step size} (device doubles; start
 * with {0, initial_step}).  Appends at most cap accepted states to
 * traj [cap][8][n]; *count (device int64) = states appended; *flag (device
 * int32) = BH_DENSE_DONE (elapsed >= lambda_total), BH_DENSE_CHUNK_FULL
 * (call again to continue) or BH_DENSE_UNDERFLOW (h < 1e-14: the
 * reference raises RuntimeError).  work: 24 ld + 8192 device doubles.  One
 * cooperative launch over the chip (a thread per ray up to the resident
 * capacity, a grid-wide barrier per attempt); the trajectory does not
 * depend on the grid size. */
#define BH_DENSE_DONE 0
#define BH_DENSE_CHUNK_FULL 1
#define BH_DENSE_UNDERFLOW 2
int32_t bh_integrate_dense_f64(double mass, double spin, int32_t mode, double rtol,
                               double atol, double max_step, double lambda_total,
                               double *state_soa, int64_t ld, int64_t n, double *work,
                               double *traj, int64_t cap, double *ctl, int64_t *count,
                               int32_t *flag, void *stream);

/* PinholeCamera rays built on the device (camera.py:96-120, then
 * null_momentum_from_velocity and build_state, geodesics.py:25-39 and
 * 109-170): camera = {position[3], forward[3], right[3], up[3]} (host
 * doubles), half_w = tan(fov / 2), half_h = half_w * height / width; pixel
 * i = row * width + col gets state (0, position, p) in state_soa [8][ld];
 * past != 0 for past-directed photons. */
int32_t bh_camera_rays_f64(double mass, double spin, const double *camera, double half_w,
                           double half_h, int64_t width, int64_t height, int32_t past,
                           double *state_soa, int64_t ld, void *stream);

/* celestial_checkerboard (imaging.py:27-50): RGB uint8 (n, 3) of n unit
 * directions (n, 3) on `cells` polar cells (cells >= 1). */
int32_t bh_checker_colors_u8(const double *directions, int64_t n, int32_t cells, uint8_t *rgb,
                             void *stream);

/* render_image (imaging.py:53-91) on the device: RGB uint8 (n, 3) per ray
 * from final states (SoA) and status -- captured (0,0,0), escaped by the
 * celestial checkerboard of the final coordinate-velocity direction with
 * `cells` polar cells (imaging.py:27-50), MAX_STEPS (120,32,32), any other
 * status (200,0,200). */
int32_t bh_shade_checker_u8(double mass, double spin, const double *state_soa, int64_t ld,
                            const int32_t *status, int64_t n, int32_t cells, uint8_t *rgb,
                            void *stream);

/* null_momentum_from_velocity (geodesics.py:109-170): positions (n, 3),
 * directions (n, 3) -> covariant momenta (n, 4).  past != 0 picks the
 * past-directed root; normalize != 0 rescales to |p_t| = 1. */
int32_t bh_null_momentum_f64(double mass, double spin, const double *positions,
                             const double *directions, int64_t n, int32_t past,
                             int32_t normalize, double *momenta, void *stream);

/* timelike_momentum (geodesics.py:173-209): positions (n, 3), dx^i/dt (n, 3)
 * -> momenta (n, 4); *bad_count (device int32) receives the number of
 * non-timelike inputs (the host raises ValueError when nonzero). */
int32_t bh_timelike_momentum_f64(double mass, double spin,
                                 const double *positions,
                                 const double *velocities, int64_t n,
                                 double *momenta, int32_t *bad_count,
                                 void *stream);

/* hamiltonian (geodesics.py:77-85), conserved_quantities (88-92) and
 * coordinate_velocity (95-106) of (n, 8) rows.  Any output may be NULL:
 * ham (n), energy (n), l_z (n), velocity (n, 4). */
int32_t bh_diagnostics_f64(double mass, double spin, const double *state,
                           int64_t n, double *ham, double *energy, double *l_z,
                           double *velocity, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* BH_TRACE_H */
