// Internal declarations shared by the two arithmetic builds of the kernels
// (bh_precise.hip, bh_fast.hip) and the C-ABI layer (bh_abi.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bh_trace.h"

namespace bh {

// Kernel-side copy of one job's controller settings, with the spin-derived
// constants the geometry uses.  Each derived constant is the same double the
// reference forms from Python floats (kerr.py:134,138,165,166), so the
// per-ray arithmetic sees identical operands.
struct JobConst {
  double m, a, a2;
  double four_a2;    // 4.0 * a2         (kerr.py:138)
  double three_a2;   // 3.0 * a2         (kerr.py:165)
  double dhz_c;      // -2.0 * m * a2    (kerr.py:166)
  double escape_radius, capture_radius;
  double rtol, atol;
  double initial_step, max_step, min_step;
  int64_t max_steps;
  int64_t max_attempts;  // 4 * max_steps  (tracer.py:101)
  double lambda_end;     // +inf when the caller sets none
};

__host__ __device__ inline JobConst make_job_const(const bh_trace_params &p) {
  JobConst c;
  c.m = p.mass;
  c.a = p.spin;
  c.a2 = p.spin * p.spin;
  c.four_a2 = 4.0 * c.a2;
  c.three_a2 = 3.0 * c.a2;
  c.dhz_c = -2.0 * p.mass * c.a2;
  c.escape_radius = p.escape_radius;
  c.capture_radius = p.capture_radius;
  c.rtol = p.rtol;
  c.atol = p.atol;
  c.initial_step = p.initial_step;
  c.max_step = p.max_step;
  c.min_step = p.min_step;
  c.max_steps = p.max_steps;
  c.max_attempts = 4 * p.max_steps;
  c.lambda_end = p.lambda_end > 0.0 ? p.lambda_end : __builtin_inf();  // <= 0 or NaN: none
  return c;
}

struct TraceIO {
  const double *in;
  int64_t ld_in;
  double *out;
  int64_t ld_out;
  int32_t *status;
  int64_t *steps;
  int64_t *attempts;
  unsigned long long *iterations;
  const int64_t *order;  // optional: k-th ray handed out is order[k]
  int64_t n;
};

// Values per checkpoint: state (8), controller step size, affine parameter.
constexpr int CKPT_VALUES = 10;

// Forward-with-checkpoints outputs (SoA, plane stride n).
struct RecordIO {
  double *ckpt;      // [cap][CKPT_VALUES][n] every K accepted steps
  int32_t *n_ckpt;   // checkpoints written per ray
  double *term;      // [8][n]: terminal point
  double *prev;      // [9][n]: start state and step size of the event step
  double *evy1;      // [8][n]: state at the end of the event step
  int64_t *ev_step;  // [n]: accepted-step index at the end of the event step
  double *theta;     // [n]: dense-output parameter of the event in that step
  int32_t *event;    // [n]: 0 none, 1 escape sphere, 2 scoring sphere, -1 checkpoint overflow
  // optional trajectory store (instead of checkpoints): pages of
  // TRAJ_PAGE_STEPS accepted steps taken from a pool by an atomic counter;
  // step j of page p (start state and step size, TRAJ_VALUES doubles) at
  // pool[(p * TRAJ_PAGE_STEPS + j) * TRAJ_VALUES]; page_link[p] is the
  // ray's previous page (-1 for its first), last_page[i] ray i's last page;
  // last_page != NULL selects the trajectory store
  double *pool;
  int64_t *page_link;
  int64_t *last_page;
  unsigned long long *pages_taken;  // pool counter (pages requested)
  int64_t pool_pages;
  double score_radius;
  int32_t cap;
  int32_t K;
};

// Values per stored trajectory step: the step's start state (8) and its
// accepted step size.
constexpr int TRAJ_VALUES = 9;
constexpr int TRAJ_PAGE_STEPS = BH_TRAJ_PAGE_STEPS;
constexpr int TRAJ_RUN_MAX = BH_TRAJ_RUN_MAX;
static_assert((TRAJ_PAGE_STEPS & (TRAJ_PAGE_STEPS - 1)) == 0, "page size: a power of two");

struct AdjointIO {
  const double *ckpt;
  const int32_t *event;
  const int64_t *ev_step;
  const double *evy1;
  const double *prev;
  const double *theta;
  const double *gterm;  // [8][n] adjoint of the terminal point
  double *gy0;          // [8][n] adjoint of the initial state
  double *gspin;        // [n] per-ray spin adjoint
  double *scratch;      // [ceil(n / 64)][9 K][64]
  const int64_t *order; // optional [n]: lane slot -> ray
  // optional trajectory store (RecordIO's pool, page_link, last_page) and
  // each ray's accepted-step count: no replay
  const double *pool;
  const int64_t *page_link;
  const int64_t *last_page;
  const int64_t *steps;
  int64_t pool_pages;
  double score_radius;
  int64_t n;
  int32_t K;
};

// Pinhole camera: position, orthonormal basis and half extents.
struct CameraConst {
  double pos[3], fwd[3], right[3], up[3];
  double half_w, half_h;
};

// Offline tile tracer (offline/render.py:184-349).
struct TileConst {
  double disk_inner, disk_outer;
  double interior_near;    // capture_radius * 1.002 (render.py:243-246)
  int64_t saturate_steps;  // int(0.85 * max_steps) (render.py:294-296)
  int32_t disk, interior;
};

struct TileIO {
  const double *in;
  int64_t ld_in;
  double *out;  // optional final states
  int64_t ld_out;
  int32_t *status;
  int64_t *steps;
  int64_t *attempts;
  uint8_t *saturated;
  double *escape_dir;  // (n, 3)
  double *hit_pos;     // (n, 3)
  double *hit_mom;     // (n, 3)
  unsigned long long *iterations;
  const int64_t *order;
  int64_t n;
};

// The real-time shader's fixed-step algorithm (realtime/reference.py:41-188).
struct ShaderConst {
  double r_plus, step_scale, min_step, max_step, bailout2;
  double disk_inner, disk_outer, interior_near;
  int64_t max_steps;
  int32_t disk, interior;
};

struct ShaderIO {
  const double *in;
  int64_t ld_in;
  double *out;  // optional final states
  int64_t ld_out;
  int32_t *status;
  double *hit;  // optional (n, 8): crossing radius, position (3), momentum (4)
  const int64_t *order;
  int64_t n;
};

// Batch-shared dense integration (integrators.py:127-174), one chunk.
struct DenseIO {
  double *y;            // [8][ld] in: states at the chunk's start; out: at its end
  int64_t ld;
  double *work;         // 24 ld + 8192 scratch: SoA planes, then 2 planes of per-block maxima
  double *traj;         // [cap][8][n] accepted states of this chunk
  double *ctl;          // [2] elapsed affine parameter, step size (in/out)
  int64_t *count;       // [1] accepted states written
  int32_t *flag;        // [1] BH_DENSE_*
  int64_t n, cap;
  double lambda_total, max_step;
};

// Doubles of adjoint scratch per ray: K replayed steps (state and step size
// each).
__host__ __device__ constexpr int64_t adjoint_scratch_per_ray(int64_t K) { return 9 * K; }

// Jobs one batched launch can hold (job table and ray offsets live in LDS).
constexpr int MAX_BATCH_JOBS = 256;

struct LaunchShape {
  int block;
  int rays_per_lane;
  int refill_min;
};

// Host launchers implemented once per arithmetic build.
#define BH_DECLARE_LAUNCHERS(NS)                                                   \
  namespace NS {                                                                   \
  hipError_t launch_trace_plain(const JobConst &c, const TraceIO &io,              \
                                const LaunchShape &shape, hipStream_t s);          \
  hipError_t launch_trace_batched(const bh_trace_params *jobs, const int64_t *offsets,   \
                                  int n_jobs, const TraceIO &io,                   \
                                  unsigned long long *queue, const LaunchShape &shape, \
                                  hipStream_t s);                                  \
  hipError_t launch_trace_record(const JobConst &c, const TraceIO &io,            \
                                 const RecordIO &rec, hipStream_t s);              \
  hipError_t launch_trace_adjoint(const JobConst &c, const AdjointIO &a,           \
                                  hipStream_t s);                                  \
  hipError_t launch_rhs_vjp(const JobConst &c, const double *state, const double *v, \
                            int64_t n, double *zbar, double *abar, hipStream_t s); \
  hipError_t launch_radial_minimum(double mass, const double *spin, const double *lam, \
                                   const double *eta, const double *r_hi, int64_t n, int grid, \
                                   int refine, double *r_star, int32_t *at_lo, hipStream_t s); \
  hipError_t launch_integrate_dense(const JobConst &c, const DenseIO &d, hipStream_t s); \
  hipError_t kernel_attributes(int which, int block, int32_t *out);                \
  hipError_t launch_trace_shader(const JobConst &c, const ShaderConst &sp, const ShaderIO &io, \
                                 hipStream_t s);                                   \
  hipError_t launch_trace_tile(const JobConst &c, const TileConst &t, const TileIO &io, \
                               int block, hipStream_t s);                         \
  hipError_t launch_midpoint(const JobConst &c, double *y_soa, int64_t ld, const double *h, \
                             int64_t n, int64_t n_steps, int iterations, hipStream_t s); \
  hipError_t launch_camera_rays(const JobConst &c, const CameraConst &cam, int64_t width, \
                                int64_t height, int past, double *soa, int64_t ld, \
                                hipStream_t s);                                    \
  hipError_t launch_shade_checker(const JobConst &c, const double *soa, int64_t ld, \
                                  const int32_t *status, int64_t n, int cells,     \
                                  uint8_t *rgb, hipStream_t s);                     \
  hipError_t launch_rk4(const JobConst &c, double *y_soa, int64_t ld, const double *h, \
                        int64_t n, int64_t n_steps, hipStream_t s);               \
  hipError_t launch_count_outside(const JobConst &c, const double *in, int64_t ld, \
                                  int64_t n, int32_t *count, hipStream_t s);       \
  hipError_t launch_geometry(const JobConst &c, const double *xyz, int64_t n,      \
                             double *radius, double *h, double *l, double *gh,     \
                             double *gl, hipStream_t s);                           \
  hipError_t launch_rhs(const JobConst &c, const double *state, int64_t n,         \
                        double *deriv, hipStream_t s);                             \
  hipError_t launch_error_ratio(const double *state, const double *new_state,     \
                                const double *error, int64_t n, int dim,           \
                                double rtol, double atol, double *ratio,           \
                                hipStream_t s);                                    \
  hipError_t launch_step_factor(const double *ratio, int64_t n, double *factor,    \
                                hipStream_t s);                                    \
  hipError_t launch_checker_colors(const double *dirs, int64_t n, int cells,       \
                                   uint8_t *rgb, hipStream_t s);                   \
  hipError_t launch_dp5_step(const JobConst &c, const double *state,               \
                             const double *h, int64_t n, double *new_state,        \
                             double *err, double *ratio, double *factor,           \
                             hipStream_t s);                                       \
  hipError_t launch_null_momentum(const JobConst &c, const double *pos,            \
                                  const double *dir, int64_t n, int past,          \
                                  int normalize, double *mom, hipStream_t s);      \
  hipError_t launch_timelike_momentum(const JobConst &c, const double *pos,        \
                                      const double *vel, int64_t n, double *mom,   \
                                      int32_t *bad, hipStream_t s);                \
  hipError_t launch_diagnostics(const JobConst &c, const double *state, int64_t n, \
                                double *ham, double *energy, double *lz,           \
                                double *vel, hipStream_t s);                       \
  }

BH_DECLARE_LAUNCHERS(precise)
BH_DECLARE_LAUNCHERS(fast)

namespace precise {
// The batched tracer with a DPP quad of lanes per ray (precise build only).
hipError_t launch_trace_quad(const bh_trace_params *jobs, const int64_t *offsets, int n_jobs,
                             const TraceIO &io, unsigned long long *queue,
                             const LaunchShape &shape, hipStream_t s);
// The recording forward with a quad per ray, one wavefront per SIMD on
// `cus` compute units.
hipError_t launch_trace_record_quad(const JobConst &c, const TraceIO &io, const RecordIO &rec,
                                    unsigned long long *queue, int cus, hipStream_t s);
}  // namespace precise

hipError_t launch_aos_to_soa(const double *aos, double *soa, int64_t ld, int64_t n,
                             hipStream_t s);
hipError_t launch_soa_to_aos(const double *soa, int64_t ld, double *aos, int64_t n,
                             hipStream_t s);
hipError_t launch_zero_u64(unsigned long long *p, hipStream_t s);
hipError_t launch_zero_i32(int32_t *p, hipStream_t s);

}  // namespace bh
