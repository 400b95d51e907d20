// C ABI of the tracer (declared in include/bh_trace.h): argument checks,
// arithmetic-mode dispatch and the small layout/fill kernels.
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <mutex>

#include "bh_internal.h"

namespace bh {

static thread_local char g_err[512] = "";

static int32_t fail(int32_t code, const char *fmt, const char *detail = nullptr) {
  if (detail) snprintf(g_err, sizeof(g_err), fmt, detail);
  else snprintf(g_err, sizeof(g_err), "%s", fmt);
  return code;
}

static int32_t hip_status(hipError_t e, const char *where) {
  if (e == hipSuccess) return BH_OK;
  snprintf(g_err, sizeof(g_err), "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
  return BH_EHIP;
}

__global__ void aos_to_soa_kernel(const double *aos, double *soa, int64_t ld, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 8; ++c) soa[c * ld + i] = aos[8 * i + c];
}

__global__ void soa_to_aos_kernel(const double *soa, int64_t ld, double *aos, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 8; ++c) aos[8 * i + c] = soa[c * ld + i];
}

__global__ void zero_u64_kernel(unsigned long long *p) { *p = 0ull; }
__global__ void zero_i32_kernel(int32_t *p) { *p = 0; }

hipError_t launch_aos_to_soa(const double *aos, double *soa, int64_t ld, int64_t n,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(aos_to_soa_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, aos,
                     soa, ld, n);
  return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double *soa, int64_t ld, double *aos, int64_t n,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(soa_to_aos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, soa,
                     ld, aos, n);
  return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *p, hipStream_t s) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_zero_i32(int32_t *p, hipStream_t s) {
  hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, s, p);
  return hipGetLastError();
}

// Spacetime checks of KerrSpacetime.__init__ (kerr.py:68-71).
static int32_t check_spacetime(double mass, double spin) {
  if (!(mass > 0.0)) return fail(BH_EINVAL, "mass must be positive");
  if (!(fabs(spin) <= mass)) return fail(BH_EINVAL, "Kerr bound violated: |spin| must be <= mass");
  return BH_OK;
}

// Largest ray count an entry accepts: 2^36 rays keep every grid below 2^32
// workgroups down to 16-thread blocks (one SoA plane of that many rays is
// 512 GiB, beyond one GPU's memory), and ray-count arithmetic far from int64
// overflow.
constexpr int64_t MAX_RAYS = (int64_t)1 << 36;

static bool rays_ok(int64_t n) { return n >= 0 && n <= MAX_RAYS; }

// Compute units of the current device (cached per thread and device).
static hipError_t device_cus(int *cus) {
  static thread_local int cached_device = -1, cached_cus = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev != cached_device) {
    e = hipDeviceGetAttribute(&cached_cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    cached_device = dev;
  }
  *cus = cached_cus;
  return hipSuccess;
}

static int32_t check_mode(int32_t mode) {
  if (mode != BH_MODE_FAST && mode != BH_MODE_PRECISE)
    return fail(BH_EINVAL, "mode must be BH_MODE_FAST or BH_MODE_PRECISE");
  return BH_OK;
}

constexpr int64_t MAX_STEPS_ABS = (int64_t)1 << 60;

static int32_t check_params(const bh_trace_params *p) {
  if (!p) return fail(BH_EINVAL, "params is NULL");
  int32_t rc = check_spacetime(p->mass, p->spin);
  if (rc) return rc;
  if (p->mode != BH_MODE_FAST && p->mode != BH_MODE_PRECISE)
    return fail(BH_EINVAL, "mode must be BH_MODE_FAST or BH_MODE_PRECISE");
  // |max_steps| <= 2^60 keeps the attempt cap 4 * max_steps inside int64; a
  // negative max_steps traces nothing (every ray MAX_STEPS), as the
  // reference's loop bound 4 * max_steps <= 0 runs no iteration.
  if (p->max_steps > MAX_STEPS_ABS || p->max_steps < -MAX_STEPS_ABS)
    return fail(BH_EINVAL, "max_steps out of range [-2^60, 2^60]");
  return BH_OK;
}

static JobConst spacetime_const(double mass, double spin) {
  bh_trace_params p;
  memset(&p, 0, sizeof(p));
  p.mass = mass;
  p.spin = spin;
  return make_job_const(p);
}

}  // namespace bh

using namespace bh;

extern "C" {

int32_t bh_abi_version(void) { return BH_ABI_VERSION; }

const char *bh_last_error(void) { return g_err; }

int32_t bh_count_outside_f64(const bh_trace_params *p, const double *state_soa, int64_t ld,
                             int64_t n, int32_t *count_out, void *stream) {
  int32_t rc = check_params(p);
  if (rc) return rc;
  if (!rays_ok(n) || ld < n) return fail(BH_EINVAL, "bad n / ld");
  if (!count_out || (n > 0 && !state_soa)) return fail(BH_EINVAL, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  JobConst c = make_job_const(*p);
  hipError_t e = launch_zero_i32(count_out, s);
  if (e == hipSuccess)
    e = p->mode == BH_MODE_PRECISE ? precise::launch_count_outside(c, state_soa, ld, n, count_out, s)
                                   : fast::launch_count_outside(c, state_soa, ld, n, count_out, s);
  return hip_status(e, "bh_count_outside_f64");
}

int32_t bh_trace_plain_f64(const bh_trace_params *p, const bh_tune_params *tune,
                           const double *state_in_soa, int64_t ld_in, double *state_out_soa,
                           int64_t ld_out, int32_t *status, int64_t *steps, int64_t *attempts,
                           uint64_t *iterations, const int64_t *order, int64_t n,
                           void *stream) {
  int32_t rc = check_params(p);
  if (rc) return rc;
  if (!rays_ok(n) || ld_in < n || ld_out < n) return fail(BH_EINVAL, "bad n / ld");
  if (n > 0 && (!state_in_soa || !state_out_soa || !status || !steps || !attempts))
    return fail(BH_EINVAL, "NULL buffer");
  LaunchShape shape{256, 1, 1};
  if (tune) {
    if (tune->block_size) shape.block = tune->block_size;
    if (tune->rays_per_lane) shape.rays_per_lane = tune->rays_per_lane;
    if (tune->refill_min) shape.refill_min = tune->refill_min;
    if (tune->lanes_per_ray != 0 && tune->lanes_per_ray != 1)
      return fail(BH_EINVAL, "the plain tracer takes lanes_per_ray 1");
  }
  if (shape.block <= 0 || shape.block > 256 || shape.block % 64)
    return fail(BH_EINVAL, "block_size must be a multiple of 64 in [64, 256]");
  if (shape.rays_per_lane < 1) return fail(BH_EINVAL, "rays_per_lane must be >= 1");
  if (shape.refill_min < 1 || shape.refill_min > 64)
    return fail(BH_EINVAL, "refill_min must be in [1, 64]");
  hipStream_t s = (hipStream_t)stream;
  JobConst c = make_job_const(*p);
  TraceIO io{state_in_soa, ld_in, state_out_soa, ld_out, status, steps, attempts,
             (unsigned long long *)iterations, order, n};
  hipError_t e = hipSuccess;
  if (iterations) e = launch_zero_u64((unsigned long long *)iterations, s);
  if (e == hipSuccess)
    e = p->mode == BH_MODE_PRECISE ? precise::launch_trace_plain(c, io, shape, s)
                                   : fast::launch_trace_plain(c, io, shape, s);
  return hip_status(e, "bh_trace_plain_f64");
}

int32_t bh_trace_batched_f64(const bh_trace_params *jobs, const int64_t *ray_offsets,
                             int32_t n_jobs, int32_t mode, const bh_tune_params *tune,
                             const double *state_in_soa, int64_t ld_in, double *state_out_soa,
                             int64_t ld_out, int32_t *status, int64_t *steps, int64_t *attempts,
                             uint64_t *iterations, const int64_t *order, uint64_t *workspace,
                             int64_t n_total, void *stream) {
  if (n_jobs < 1 || n_jobs > MAX_BATCH_JOBS)
    return fail(BH_EINVAL, "n_jobs must be in [1, 256]");
  if (mode != BH_MODE_FAST && mode != BH_MODE_PRECISE)
    return fail(BH_EINVAL, "mode must be BH_MODE_FAST or BH_MODE_PRECISE");
  if (!rays_ok(n_total) || ld_in < n_total || ld_out < n_total) return fail(BH_EINVAL, "bad n / ld");
  if (!jobs || !ray_offsets || !iterations || !workspace) return fail(BH_EINVAL, "NULL buffer");
  if (n_total > 0 && (!state_in_soa || !state_out_soa || !status || !steps || !attempts))
    return fail(BH_EINVAL, "NULL buffer");
  LaunchShape shape{256, 1, 1};
  int lanes = 1;
  if (tune) {
    if (tune->block_size) shape.block = tune->block_size;
    if (tune->refill_min) shape.refill_min = tune->refill_min;
    if (tune->lanes_per_ray) lanes = tune->lanes_per_ray;
  }
  if (shape.block <= 0 || shape.block > 256 || shape.block % 64)
    return fail(BH_EINVAL, "block_size must be a multiple of 64 in [64, 256]");
  if (lanes != 1 && lanes != 4) return fail(BH_EINVAL, "lanes_per_ray must be 1 or 4");
  if (lanes == 4 && mode != BH_MODE_PRECISE)
    return fail(BH_EINVAL, "lanes_per_ray 4 needs the precise build");
  if (shape.refill_min < 1 || shape.refill_min > 64 / lanes)
    return fail(BH_EINVAL, lanes == 4 ? "refill_min must be in [1, 16] with lanes_per_ray 4"
                                      : "refill_min must be in [1, 64]");
  TraceIO io{state_in_soa, ld_in, state_out_soa, ld_out, status, steps, attempts,
             (unsigned long long *)iterations, order, n_total};
  hipStream_t s = (hipStream_t)stream;
  unsigned long long *queue = (unsigned long long *)workspace;
  hipError_t e =
      lanes == 4 ? precise::launch_trace_quad(jobs, ray_offsets, n_jobs, io, queue, shape, s)
      : mode == BH_MODE_PRECISE
          ? precise::launch_trace_batched(jobs, ray_offsets, n_jobs, io, queue, shape, s)
          : fast::launch_trace_batched(jobs, ray_offsets, n_jobs, io, queue, shape, s);
  return hip_status(e, "bh_trace_batched_f64");
}

// Checkpoint intervals beyond this are refused (the per-ray scratch count
// must fit int32).
constexpr int32_t MAX_ADJOINT_K = 1 << 20;

// Argument checks shared by the two reverse-pass entries.
static int32_t check_adjoint_args(const bh_trace_params *p, int64_t n, int32_t K,
                                  const double *ckpt, const int32_t *event,
                                  const int64_t *ev_step, const double *evy1, const double *prev,
                                  const double *theta, const double *gterm, double *gy0,
                                  double *gspin) {
  int32_t rc = check_params(p);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (K < 1 || K > MAX_ADJOINT_K) return fail(BH_EINVAL, "need 1 <= K <= 2^20");
  if (n > 0 && (!ckpt || !event || !ev_step || !evy1 || !prev || !theta || !gterm || !gy0 ||
                !gspin))
    return fail(BH_EINVAL, "NULL buffer");
  return BH_OK;
}

// The trajectory store a reverse pass reads (last_page NULL: checkpoints).
struct TrajIn {
  const double *pool;
  const int64_t *page_link, *last_page, *steps;
  int64_t pool_pages;
};

static int32_t launch_adjoint(const bh_trace_params *p, int64_t n, int32_t K, double score_radius,
                              const double *ckpt, const int32_t *event, const int64_t *ev_step,
                              const double *evy1, const double *prev, const double *theta,
                              const double *gterm, double *gy0, double *gspin, double *scratch,
                              const int64_t *order, const TrajIn &traj, void *stream,
                              const char *name) {
  JobConst c = make_job_const(*p);
  AdjointIO a{ckpt,    event,    ev_step,    evy1,  prev,       theta,        gterm,
              gy0,     gspin,    scratch,    order, traj.pool, traj.page_link,
              traj.last_page, traj.steps, traj.pool_pages, score_radius, n, K};
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = p->mode == BH_MODE_PRECISE ? precise::launch_trace_adjoint(c, a, s)
                                            : fast::launch_trace_adjoint(c, a, s);
  return hip_status(e, name);
}

// Scratch of bh_trace_adjoint_static_f64: one copy of this array lives on
// every GPU the library runs on, from the first use there until the process
// ends.  In host code its name is only a stand-in for the symbol, never a
// device address: the device address is resolved once per device with
// hipGetSymbolAddress (that device current) and cached, and only that
// address reaches kernels.  Calls that overlap on one GPU share the array
// and race; the backward writes every slot it reads in the same call, so
// nothing needs clearing between calls.
static __device__ __align__(256) double scratch[BH_ADJOINT_STATIC_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[MAX_DEVICES];

static int32_t static_scratch(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "bh_trace_adjoint_static_f64: hipGetDevice");
  if (dev < 0 || dev >= MAX_DEVICES)
    return fail(BH_EINVAL, "bh_trace_adjoint_static_f64: device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return hip_status(e, "bh_trace_adjoint_static_f64: hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(BH_EHIP, "bh_trace_adjoint_static_f64: scratch has no device address");
    g_scratch_ptr[dev] = (double *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return BH_OK;
}

int32_t bh_adjoint_scratch_doubles(int32_t K) {
  return K < 1 || K > MAX_ADJOINT_K ? -1 : (int32_t)adjoint_scratch_per_ray(K);
}

// The recording launch shared by the checkpoint and trajectory entries.
static int32_t launch_record(const bh_trace_params *p, const TraceIO &io, const RecordIO &rec,
                             uint64_t *workspace, void *stream, const char *name) {
  const int64_t n = io.n;
  unsigned long long *iterations = io.iterations;
  hipStream_t s = (hipStream_t)stream;
  JobConst c = make_job_const(*p);
  hipError_t e = hipSuccess;
  if (iterations) e = launch_zero_u64((unsigned long long *)iterations, s);
  if (e == hipSuccess && rec.last_page) e = launch_zero_u64(rec.pages_taken, s);
  int cus = 0;
  if (e == hipSuccess && n > 0 && p->mode == BH_MODE_PRECISE) e = device_cus(&cus);
  if (e == hipSuccess && n > 0) {
    // precise recordings of at most one wavefront per SIMD are as long as
    // their longest rays: a quad per ray shortens each attempt (same bits)
    if (p->mode == BH_MODE_PRECISE && (n + 63) / 64 <= (int64_t)cus * 4)
      e = precise::launch_trace_record_quad(c, io, rec, (unsigned long long *)workspace, cus, s);
    else
      e = p->mode == BH_MODE_PRECISE ? precise::launch_trace_record(c, io, rec, s)
                                     : fast::launch_trace_record(c, io, rec, s);
  }
  return hip_status(e, name);
}

int32_t bh_trace_record_f64(const bh_trace_params *p, const double *state_in_soa, int64_t ld_in,
                            double *state_out_soa, int64_t ld_out, int32_t *status,
                            int64_t *steps, int64_t *attempts, uint64_t *iterations,
                            const int64_t *order, int64_t n, int32_t K, int32_t cap,
                            double score_radius, double *ckpt,
                            int32_t *n_ckpt, double *term, double *prev, double *evy1,
                            int64_t *ev_step, double *theta, int32_t *event, uint64_t *workspace,
                            void *stream) {
  int32_t rc = check_params(p);
  if (rc) return rc;
  if (!rays_ok(n) || ld_in < n || ld_out < n) return fail(BH_EINVAL, "bad n / ld");
  if (K < 1 || K > MAX_ADJOINT_K || cap < 1)
    return fail(BH_EINVAL, "need 1 <= K <= 2^20 and cap >= 1");
  if (n > 0 && (!state_in_soa || !state_out_soa || !status || !steps || !attempts || !ckpt ||
                !n_ckpt || !term || !prev || !evy1 || !ev_step || !theta || !event || !workspace))
    return fail(BH_EINVAL, "NULL buffer");
  TraceIO io{state_in_soa, ld_in, state_out_soa, ld_out, status, steps, attempts,
             (unsigned long long *)iterations, order, n};
  RecordIO rec{ckpt,  n_ckpt,  term,    prev,    evy1,         ev_step, theta, event,
               nullptr, nullptr, nullptr, nullptr, 0,       score_radius, cap,   K};
  return launch_record(p, io, rec, workspace, stream, "bh_trace_record_f64");
}

// Pool pages beyond this are refused (page * BH_TRAJ_PAGE_STEPS * 9 must
// fit int64 with room to spare).
constexpr int64_t MAX_POOL_PAGES = (int64_t)1 << 48;

int32_t bh_trace_record_trajectory_f64(
    const bh_trace_params *p, const double *state_in_soa, int64_t ld_in, double *state_out_soa,
    int64_t ld_out, int32_t *status, int64_t *steps, int64_t *attempts, uint64_t *iterations,
    const int64_t *order, int64_t n, double *pool, int64_t pool_pages, int64_t *page_link,
    int64_t *last_page, double score_radius, double *term, double *prev, double *evy1,
    int64_t *ev_step, double *theta, int32_t *event, uint64_t *workspace, void *stream) {
  int32_t rc = check_params(p);
  if (rc) return rc;
  if (!rays_ok(n) || ld_in < n || ld_out < n) return fail(BH_EINVAL, "bad n / ld");
  if (pool_pages < 0 || pool_pages > MAX_POOL_PAGES)
    return fail(BH_EINVAL, "pool_pages out of range [0, 2^48]");
  if (n > 0 && (!state_in_soa || !state_out_soa || !status || !steps || !attempts ||
                !last_page || !term || !prev || !evy1 || !ev_step || !theta || !event ||
                !workspace || (pool_pages > 0 && (!pool || !page_link))))
    return fail(BH_EINVAL, "NULL buffer");
  TraceIO io{state_in_soa, ld_in, state_out_soa, ld_out, status, steps, attempts,
             (unsigned long long *)iterations, order, n};
  // last_page marks the recording as a trajectory store (with a pool of
  // zero pages every ray with a step overflows)
  RecordIO rec{nullptr, nullptr,    term,      prev,      evy1,
               ev_step, theta,      event,     pool,      page_link,
               last_page, (unsigned long long *)workspace + 1, pool_pages, score_radius, 0,
               1};
  return launch_record(p, io, rec, workspace, stream, "bh_trace_record_trajectory_f64");
}


int32_t bh_trace_adjoint_f64(const bh_trace_params *p, int64_t n, int32_t K, double score_radius,
                             const double *ckpt, const int32_t *event, const int64_t *ev_step,
                             const double *evy1, const double *prev, const double *theta,
                             const double *gterm, double *gy0, double *gspin, double *scratch,
                             const int64_t *order, void *stream) {
  int32_t rc = check_adjoint_args(p, n, K, ckpt, event, ev_step, evy1, prev, theta, gterm, gy0,
                                  gspin);
  if (rc) return rc;
  if (n > 0 && !scratch) return fail(BH_EINVAL, "NULL buffer");
  return launch_adjoint(p, n, K, score_radius, ckpt, event, ev_step, evy1, prev, theta, gterm,
                        gy0, gspin, scratch, order, TrajIn{}, stream,
                        "bh_trace_adjoint_f64");
}

int32_t bh_trace_adjoint_static_f64(const bh_trace_params *p, int64_t n, int32_t K,
                                    double score_radius, const double *ckpt, const int32_t *event,
                                    const int64_t *ev_step, const double *evy1,
                                    const double *prev, const double *theta, const double *gterm,
                                    double *gy0, double *gspin, const int64_t *order,
                                    void *stream) {
  int32_t rc = check_adjoint_args(p, n, K, ckpt, event, ev_step, evy1, prev, theta, gterm, gy0,
                                  gspin);
  if (rc) return rc;
  if (n == 0) return BH_OK;
  // per wavefront of 64 rays, 9 K doubles per ray (n <= 2^36, K <= 2^20: no overflow)
  const int64_t need = adjoint_scratch_per_ray(K) * ((n + 63) / 64) * 64;
  if (need > BH_ADJOINT_STATIC_SCRATCH_DOUBLES)
    return fail(BH_EINVAL,
                "bh_trace_adjoint_static_f64: 9 K doubles per ray (n rounded up to a multiple "
                "of 64) exceed the static scratch of BH_ADJOINT_STATIC_SCRATCH_DOUBLES (2^23); "
                "use bh_trace_adjoint_f64 with a caller scratch buffer");
  double *ptr = nullptr;
  rc = static_scratch(&ptr);
  if (rc) return rc;
  return launch_adjoint(p, n, K, score_radius, ckpt, event, ev_step, evy1, prev, theta, gterm,
                        gy0, gspin, ptr, order, TrajIn{}, stream,
                        "bh_trace_adjoint_static_f64");
}

int32_t bh_trace_adjoint_trajectory_f64(const bh_trace_params *p, int64_t n,
                                        double score_radius, const double *pool,
                                        int64_t pool_pages, const int64_t *page_link,
                                        const int64_t *last_page, const int64_t *steps,
                                        const int32_t *event, const int64_t *ev_step,
                                        const double *evy1, const double *prev,
                                        const double *theta, const double *gterm, double *gy0,
                                        double *gspin, const int64_t *order, void *stream) {
  int32_t rc = check_params(p);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (pool_pages < 0 || pool_pages > MAX_POOL_PAGES)
    return fail(BH_EINVAL, "pool_pages out of range [0, 2^48]");
  if (n > 0 && (!last_page || !steps || !event || !ev_step || !evy1 || !prev || !theta ||
                !gterm || !gy0 || !gspin || (pool_pages > 0 && (!pool || !page_link))))
    return fail(BH_EINVAL, "NULL buffer");
  const TrajIn traj{pool, page_link, last_page, steps, pool_pages};
  return launch_adjoint(p, n, 1, score_radius, nullptr, event, ev_step, evy1, prev, theta, gterm,
                        gy0, gspin, nullptr, order, traj, stream,
                        "bh_trace_adjoint_trajectory_f64");
}

int32_t bh_rhs_vjp_f64(double mass, double spin, const double *state, const double *v,
                       int64_t n, double *zbar, double *abar, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && (!state || !v || !zbar || !abar)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  return hip_status(fast::launch_rhs_vjp(c, state, v, n, zbar, abar, (hipStream_t)stream),
                    "bh_rhs_vjp_f64");
}

int32_t bh_radial_minimum_f64(double mass, const double *spin, const double *lam,
                              const double *eta, const double *r_hi, int64_t n, int32_t grid,
                              int32_t refine, double *r_star, int32_t *at_lo, void *stream) {
  if (!(mass > 0.0) || !(mass < INFINITY)) return fail(BH_EINVAL, "mass must be positive");
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (grid < 2 || grid > (1 << 20) || refine < 0 || refine > 1000)
    return fail(BH_EINVAL, "need 2 <= grid <= 2^20 and 0 <= refine <= 1000");
  if (n > 0 && (!spin || !lam || !eta || !r_hi || !r_star || !at_lo))
    return fail(BH_EINVAL, "NULL buffer");
  return hip_status(fast::launch_radial_minimum(mass, spin, lam, eta, r_hi, n, grid, refine,
                                                r_star, at_lo, (hipStream_t)stream),
                    "bh_radial_minimum_f64");
}

int32_t bh_aos_to_soa_f64(const double *aos, double *soa, int64_t ld, int64_t n, void *stream) {
  if (!rays_ok(n) || ld < n) return fail(BH_EINVAL, "bad n / ld");
  if (n > 0 && (!aos || !soa)) return fail(BH_EINVAL, "NULL buffer");
  return hip_status(launch_aos_to_soa(aos, soa, ld, n, (hipStream_t)stream), "bh_aos_to_soa_f64");
}

int32_t bh_soa_to_aos_f64(const double *soa, int64_t ld, double *aos, int64_t n, void *stream) {
  if (!rays_ok(n) || ld < n) return fail(BH_EINVAL, "bad n / ld");
  if (n > 0 && (!aos || !soa)) return fail(BH_EINVAL, "NULL buffer");
  return hip_status(launch_soa_to_aos(soa, ld, aos, n, (hipStream_t)stream), "bh_soa_to_aos_f64");
}

int32_t bh_geometry_f64(double mass, double spin, int32_t mode, const double *xyz, int64_t n,
                        double *radius, double *h, double *l, double *grad_h, double *grad_l,
                        void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if ((rc = check_mode(mode))) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && !xyz) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = mode == BH_MODE_PRECISE
                     ? precise::launch_geometry(c, xyz, n, radius, h, l, grad_h, grad_l, s)
                     : fast::launch_geometry(c, xyz, n, radius, h, l, grad_h, grad_l, s);
  return hip_status(e, "bh_geometry_f64");
}

int32_t bh_geodesic_rhs_f64(double mass, double spin, int32_t mode, const double *state,
                            int64_t n, double *deriv, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if ((rc = check_mode(mode))) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && (!state || !deriv)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = mode == BH_MODE_PRECISE ? precise::launch_rhs(c, state, n, deriv, s)
                                         : fast::launch_rhs(c, state, n, deriv, s);
  return hip_status(e, "bh_geodesic_rhs_f64");
}

int32_t bh_error_ratio_f64(int32_t mode, const double *state, const double *new_state,
                           const double *error, int64_t n, int32_t dim, double rtol,
                           double atol, double *ratio, void *stream) {
  int32_t rc = check_mode(mode);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (dim < 1 || dim > 128) return fail(BH_EINVAL, "dim must be 1..128");
  if (n > 0 && (!state || !new_state || !error || !ratio)) return fail(BH_EINVAL, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e =
      mode == BH_MODE_PRECISE
          ? precise::launch_error_ratio(state, new_state, error, n, dim, rtol, atol, ratio, s)
          : fast::launch_error_ratio(state, new_state, error, n, dim, rtol, atol, ratio, s);
  return hip_status(e, "bh_error_ratio_f64");
}

int32_t bh_step_factor_f64(int32_t mode, const double *ratio, int64_t n, double *factor,
                           void *stream) {
  int32_t rc = check_mode(mode);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && (!ratio || !factor)) return fail(BH_EINVAL, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = mode == BH_MODE_PRECISE ? precise::launch_step_factor(ratio, n, factor, s)
                                         : fast::launch_step_factor(ratio, n, factor, s);
  return hip_status(e, "bh_step_factor_f64");
}

int32_t bh_checker_colors_u8(const double *directions, int64_t n, int32_t cells, uint8_t *rgb,
                             void *stream) {
  if (!rays_ok(n) || cells < 1) return fail(BH_EINVAL, "bad n / cells");
  if (n > 0 && (!directions || !rgb)) return fail(BH_EINVAL, "NULL buffer");
  return hip_status(precise::launch_checker_colors(directions, n, cells, rgb, (hipStream_t)stream),
                    "bh_checker_colors_u8");
}

int32_t bh_dp5_step_f64(double mass, double spin, int32_t mode, double rtol, double atol,
                        const double *state, const double *h, int64_t n, double *new_state,
                        double *error, double *ratio, double *factor, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if ((rc = check_mode(mode))) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && (!state || !h || !new_state)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  c.rtol = rtol;
  c.atol = atol;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e =
      mode == BH_MODE_PRECISE
          ? precise::launch_dp5_step(c, state, h, n, new_state, error, ratio, factor, s)
          : fast::launch_dp5_step(c, state, h, n, new_state, error, ratio, factor, s);
  return hip_status(e, "bh_dp5_step_f64");
}

int32_t bh_kernel_attributes(int32_t kernel, int32_t mode, int32_t block, int32_t *out) {
  if (!out) return fail(BH_EINVAL, "out is NULL");
  if (mode != BH_MODE_FAST && mode != BH_MODE_PRECISE) return fail(BH_EINVAL, "bad mode");
  if (kernel < BH_KERNEL_PLAIN || kernel > BH_KERNEL_QUAD) return fail(BH_EINVAL, "bad kernel");
  if (kernel == BH_KERNEL_QUAD && mode != BH_MODE_PRECISE)
    return fail(BH_EINVAL, "the quad tracer exists in the precise build only");
  if (block <= 0 || block > 256 || block % 64) return fail(BH_EINVAL, "bad block");
  if (kernel == BH_KERNEL_ADJOINT && block != 64) return fail(BH_EINVAL, "the adjoint runs 64-thread blocks");
  hipError_t e = mode == BH_MODE_PRECISE ? precise::kernel_attributes(kernel, block, out)
                                         : fast::kernel_attributes(kernel, block, out);
  return hip_status(e, "bh_kernel_attributes");
}

int32_t bh_trace_shader_f64(const bh_shader_params *p, const double *state_in_soa, int64_t ld_in,
                            double *state_out_soa, int64_t ld_out, int32_t *status, double *hit,
                            const int64_t *order, int64_t n, void *stream) {
  if (!p) return fail(BH_EINVAL, "params is NULL");
  int32_t rc = check_spacetime(p->mass, p->spin);
  if (rc) return rc;
  if (p->mode != BH_MODE_FAST && p->mode != BH_MODE_PRECISE) return fail(BH_EINVAL, "bad mode");
  if (p->max_steps < 0 || p->max_steps > ((int64_t)1 << 40)) return fail(BH_EINVAL, "bad max_steps");
  if (p->flags & ~(BH_TILE_DISK | BH_TILE_INTERIOR)) return fail(BH_EINVAL, "bad flags");
  if (!rays_ok(n) || ld_in < n || (state_out_soa && ld_out < n)) return fail(BH_EINVAL, "bad n / ld");
  if (n > 0 && (!state_in_soa || !status)) return fail(BH_EINVAL, "NULL buffer");
  bh_trace_params tp;
  memset(&tp, 0, sizeof(tp));
  tp.mass = p->mass;
  tp.spin = p->spin;
  tp.capture_radius = p->capture_radius;
  tp.escape_radius = p->escape_radius;
  JobConst c = make_job_const(tp);
  const double r_plus = p->mass + std::sqrt(p->mass * p->mass - p->spin * p->spin);
  ShaderConst sp{r_plus,        p->step_scale, p->min_step, p->max_step,
                 p->momentum_bailout * p->momentum_bailout,
                 p->disk_inner, p->disk_outer, p->capture_radius * 1.002,
                 p->max_steps,  (p->flags & BH_TILE_DISK) ? 1 : 0,
                 (p->flags & BH_TILE_INTERIOR) ? 1 : 0};
  ShaderIO io{state_in_soa, ld_in, state_out_soa, ld_out, status, hit, order, n};
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = p->mode == BH_MODE_PRECISE ? precise::launch_trace_shader(c, sp, io, s)
                                            : fast::launch_trace_shader(c, sp, io, s);
  return hip_status(e, "bh_trace_shader_f64");
}

int32_t bh_midpoint_f64(double mass, double spin, int32_t mode, double *state_soa, int64_t ld,
                        const double *h, int64_t n, int64_t n_steps, int32_t iterations,
                        void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (mode != BH_MODE_FAST && mode != BH_MODE_PRECISE) return fail(BH_EINVAL, "bad mode");
  if (!rays_ok(n) || ld < n || n_steps < 0 || iterations < 0)
    return fail(BH_EINVAL, "bad n / ld / n_steps / iterations");
  if (n > 0 && (!state_soa || !h)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e =
      mode == BH_MODE_PRECISE
          ? precise::launch_midpoint(c, state_soa, ld, h, n, n_steps, iterations, s)
          : fast::launch_midpoint(c, state_soa, ld, h, n, n_steps, iterations, s);
  return hip_status(e, "bh_midpoint_f64");
}

int32_t bh_trace_tile_f64(const bh_trace_params *p, const bh_tile_params *t,
                          const bh_tune_params *tune, const double *state_in_soa,
                          int64_t ld_in, double *state_out_soa, int64_t ld_out,
                          int32_t *status, int64_t *steps, int64_t *attempts,
                          uint8_t *saturated, double *escape_dir, double *hit_pos,
                          double *hit_mom, uint64_t *iterations, const int64_t *order,
                          int64_t n, void *stream) {
  int32_t rc = check_params(p);
  if (rc) return rc;
  if (!t) return fail(BH_EINVAL, "NULL tile parameters");
  if (!rays_ok(n) || ld_in < n || (state_out_soa && ld_out < n)) return fail(BH_EINVAL, "bad n / ld");
  if (n > 0 && (!state_in_soa || !status || !steps || !attempts || !saturated || !escape_dir ||
                !hit_pos || !hit_mom))
    return fail(BH_EINVAL, "NULL buffer");
  if (t->flags & ~(BH_TILE_DISK | BH_TILE_INTERIOR)) return fail(BH_EINVAL, "bad tile flags");
  if ((t->flags & BH_TILE_DISK) && !(t->disk_inner <= t->disk_outer))
    return fail(BH_EINVAL, "disk_inner must not exceed disk_outer");
  int block = tune && tune->block_size ? tune->block_size : 256;
  if (block <= 0 || block > 256 || block % 64)
    return fail(BH_EINVAL, "block_size must be a multiple of 64 in [64, 256]");
  JobConst c = make_job_const(*p);
  TileConst tc{t->disk_inner, t->disk_outer, p->capture_radius * 1.002, t->saturate_steps,
               (t->flags & BH_TILE_DISK) ? 1 : 0, (t->flags & BH_TILE_INTERIOR) ? 1 : 0};
  TileIO io{state_in_soa, ld_in,   state_out_soa, ld_out,  status,
            steps,        attempts, saturated,    escape_dir, hit_pos,
            hit_mom,      (unsigned long long *)iterations, order, n};
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (iterations) e = launch_zero_u64((unsigned long long *)iterations, s);
  if (e == hipSuccess)
    e = p->mode == BH_MODE_PRECISE ? precise::launch_trace_tile(c, tc, io, block, s)
                                   : fast::launch_trace_tile(c, tc, io, block, s);
  return hip_status(e, "bh_trace_tile_f64");
}

int32_t bh_rk4_f64(double mass, double spin, int32_t mode, double *state_soa, int64_t ld,
                   const double *h, int64_t n, int64_t n_steps, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (mode != BH_MODE_FAST && mode != BH_MODE_PRECISE) return fail(BH_EINVAL, "bad mode");
  if (!rays_ok(n) || ld < n || n_steps < 0) return fail(BH_EINVAL, "bad n / ld / n_steps");
  if (n > 0 && (!state_soa || !h)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = mode == BH_MODE_PRECISE ? precise::launch_rk4(c, state_soa, ld, h, n, n_steps, s)
                                         : fast::launch_rk4(c, state_soa, ld, h, n, n_steps, s);
  return hip_status(e, "bh_rk4_f64");
}

int32_t bh_integrate_dense_f64(double mass, double spin, int32_t mode, double rtol,
                               double atol, double max_step, double lambda_total,
                               double *state_soa, int64_t ld, int64_t n, double *work,
                               double *traj, int64_t cap, double *ctl, int64_t *count,
                               int32_t *flag, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (mode != BH_MODE_FAST && mode != BH_MODE_PRECISE) return fail(BH_EINVAL, "bad mode");
  if (n < 1 || !rays_ok(n) || ld < n || cap < 1) return fail(BH_EINVAL, "need n >= 1, ld >= n, cap >= 1");
  if (!(rtol >= 0.0) || !(atol >= 0.0) || !(max_step > 0.0) || std::isnan(lambda_total))
    return fail(BH_EINVAL, "bad tolerances, max_step or lambda_total");
  if (!state_soa || !work || !traj || !ctl || !count || !flag)
    return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  c.rtol = rtol;
  c.atol = atol;
  DenseIO d{state_soa, ld, work, traj, ctl, count, flag, n, cap, lambda_total, max_step};
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = mode == BH_MODE_PRECISE ? precise::launch_integrate_dense(c, d, s)
                                         : fast::launch_integrate_dense(c, d, s);
  return hip_status(e, "bh_integrate_dense_f64");
}

int32_t bh_camera_rays_f64(double mass, double spin, const double *camera, double half_w,
                           double half_h, int64_t width, int64_t height, int32_t past,
                           double *state_soa, int64_t ld, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (!camera) return fail(BH_EINVAL, "camera is NULL");
  if (width < 0 || height < 0 || (height > 0 && width > MAX_RAYS / height))
    return fail(BH_EINVAL, "bad width / height");
  if (ld < width * height) return fail(BH_EINVAL, "ld < width * height");
  if (width * height > 0 && !state_soa) return fail(BH_EINVAL, "NULL buffer");
  CameraConst cam;
  for (int j = 0; j < 3; ++j) {
    cam.pos[j] = camera[j];
    cam.fwd[j] = camera[3 + j];
    cam.right[j] = camera[6 + j];
    cam.up[j] = camera[9 + j];
  }
  cam.half_w = half_w;
  cam.half_h = half_h;
  JobConst c = spacetime_const(mass, spin);
  return hip_status(precise::launch_camera_rays(c, cam, width, height, past ? 1 : 0, state_soa,
                                                ld, (hipStream_t)stream),
                    "bh_camera_rays_f64");
}

int32_t bh_shade_checker_u8(double mass, double spin, const double *state_soa, int64_t ld,
                            const int32_t *status, int64_t n, int32_t cells, uint8_t *rgb,
                            void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (!rays_ok(n) || ld < n || cells < 1) return fail(BH_EINVAL, "bad n / ld / cells");
  if (n > 0 && (!state_soa || !status || !rgb)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  return hip_status(precise::launch_shade_checker(c, state_soa, ld, status, n, cells, rgb,
                                                  (hipStream_t)stream),
                    "bh_shade_checker_u8");
}

int32_t bh_null_momentum_f64(double mass, double spin, const double *positions,
                             const double *directions, int64_t n, int32_t past,
                             int32_t normalize, double *momenta, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && (!positions || !directions || !momenta)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  return hip_status(precise::launch_null_momentum(c, positions, directions, n, past ? 1 : 0,
                                                  normalize ? 1 : 0, momenta, (hipStream_t)stream),
                    "bh_null_momentum_f64");
}

int32_t bh_timelike_momentum_f64(double mass, double spin, const double *positions,
                                 const double *velocities, int64_t n, double *momenta,
                                 int32_t *bad_count, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && (!positions || !velocities || !momenta)) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = bad_count ? launch_zero_i32(bad_count, s) : hipSuccess;
  if (e == hipSuccess)
    e = precise::launch_timelike_momentum(c, positions, velocities, n, momenta, bad_count, s);
  return hip_status(e, "bh_timelike_momentum_f64");
}

int32_t bh_diagnostics_f64(double mass, double spin, const double *state, int64_t n, double *ham,
                           double *energy, double *l_z, double *velocity, void *stream) {
  int32_t rc = check_spacetime(mass, spin);
  if (rc) return rc;
  if (!rays_ok(n)) return fail(BH_EINVAL, "bad n");
  if (n > 0 && !state) return fail(BH_EINVAL, "NULL buffer");
  JobConst c = spacetime_const(mass, spin);
  return hip_status(
      precise::launch_diagnostics(c, state, n, ham, energy, l_z, velocity, (hipStream_t)stream),
      "bh_diagnostics_f64");
}

}  // extern "C"
