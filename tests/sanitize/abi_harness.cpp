// Host-side checks of the tracer's C ABI (include/bh_trace.h), built with
// AddressSanitizer + UndefinedBehaviorSanitizer or with ThreadSanitizer
// around the ABI translation unit.
//
// Every call below must be refused by argument validation, before any HIP
// runtime call, so the program needs no GPU and must only run where none is
// visible: the buffer arguments are host addresses that are never meant to
// reach a kernel.  Each case checks the return code (BH_EINVAL) and the
// thread-local message of bh_last_error().  Extreme sizes (INT64_MAX rays,
// width * height beyond int64) exercise the overflow guards, which UBSan
// turns into a hard failure if any size arithmetic overflows.
//
//   abi_harness cases    one pass over every refusal case
//   abi_harness threads  8 threads refusing different calls concurrently;
//                        each must read back its own message
#include <stdio.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "bh_trace.h"

namespace {

struct Case {
  std::string name;
  std::function<int32_t()> call;
  const char *message;  // substring expected in bh_last_error()
};

double g_d[64];
int32_t g_i[64];
int64_t g_l[64];
uint64_t g_u[64];
uint8_t g_b[64];
const int64_t BIG = INT64_MAX;

bh_trace_params good_params() {
  bh_trace_params p;
  memset(&p, 0, sizeof(p));
  p.mass = 1.0;
  p.spin = 0.9;
  p.escape_radius = 100.0;
  p.capture_radius = 1.45;
  p.rtol = 1e-8;
  p.atol = 1e-10;
  p.initial_step = 0.1;
  p.max_step = 10.0;
  p.min_step = 1e-12;
  p.max_steps = 1000;
  p.mode = BH_MODE_PRECISE;
  return p;
}

bh_shader_params good_shader() {
  bh_shader_params s;
  memset(&s, 0, sizeof(s));
  s.mass = 1.0;
  s.spin = 0.5;
  s.capture_radius = 1.9;
  s.escape_radius = 60.0;
  s.step_scale = 0.05;
  s.min_step = 1e-3;
  s.max_step = 2.0;
  s.momentum_bailout = 1e4;
  s.max_steps = 800;
  s.mode = BH_MODE_FAST;
  return s;
}

bh_tune_params tune_of(int32_t block, int32_t rpl, int32_t refill) {
  bh_tune_params t;
  memset(&t, 0, sizeof(t));
  t.block_size = block;
  t.rays_per_lane = rpl;
  t.refill_min = refill;
  return t;
}

std::vector<Case> build_cases() {
  std::vector<Case> v;
  double *d = g_d;
  int32_t *i32 = g_i;
  int64_t *i64 = g_l;
  uint64_t *u64 = g_u;
  auto add = [&](std::string name, std::function<int32_t()> f, const char *msg) {
    v.push_back(Case{std::move(name), std::move(f), msg});
  };

  // --- spacetime / params checks shared by the entries ---
  add("count: params NULL", [=] { return bh_count_outside_f64(nullptr, d, 4, 4, i32, nullptr); },
      "params is NULL");
  add("count: mass 0", [=] {
    bh_trace_params p = good_params();
    p.mass = 0.0;
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "mass must be positive");
  add("count: mass NaN", [=] {
    bh_trace_params p = good_params();
    p.mass = __builtin_nan("");
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "mass must be positive");
  add("count: spin 1.01", [=] {
    bh_trace_params p = good_params();
    p.spin = 1.01;
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "Kerr bound");
  add("count: spin NaN", [=] {
    bh_trace_params p = good_params();
    p.spin = __builtin_nan("");
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "Kerr bound");
  add("count: mode 7", [=] {
    bh_trace_params p = good_params();
    p.mode = 7;
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "mode must be");
  add("count: max_steps 2^62", [=] {
    bh_trace_params p = good_params();
    p.max_steps = (int64_t)1 << 62;
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "max_steps out of range");
  add("count: max_steps -2^62", [=] {
    bh_trace_params p = good_params();
    p.max_steps = -((int64_t)1 << 62);
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "max_steps out of range");
  add("count: max_steps INT64_MIN", [=] {
    bh_trace_params p = good_params();
    p.max_steps = INT64_MIN;
    return bh_count_outside_f64(&p, d, 4, 4, i32, nullptr);
  }, "max_steps out of range");
  add("adjoint static: beyond the static scratch", [=] {
    bh_trace_params p = good_params();
    return bh_trace_adjoint_static_f64(&p, 29121, 32, 4.0, d, i32, i64, d, d, d, d, d, d,
                                       nullptr, nullptr);
  }, "static scratch");
  add("adjoint static: INT64_MAX rays", [=] {
    bh_trace_params p = good_params();
    return bh_trace_adjoint_static_f64(&p, BIG, 32, 4.0, d, i32, i64, d, d, d, d, d, d,
                                       nullptr, nullptr);
  }, "bad n");
  add("adjoint static: K 2^21", [=] {
    bh_trace_params p = good_params();
    return bh_trace_adjoint_static_f64(&p, 64, 1 << 21, 4.0, d, i32, i64, d, d, d, d, d, d,
                                       nullptr, nullptr);
  }, "K <= 2^20");
  add("adjoint static: max rays at K 2^20", [=] {
    bh_trace_params p = good_params();
    return bh_trace_adjoint_static_f64(&p, (int64_t)1 << 36, 1 << 20, 4.0, d, i32, i64, d, d, d,
                                       d, d, d, nullptr, nullptr);
  }, "static scratch");
  add("plain: max_steps -2^61", [=] {
    bh_trace_params p = good_params();
    p.max_steps = -((int64_t)1 << 61);
    return bh_trace_plain_f64(&p, nullptr, d, 4, d, 4, i32, i64, i64, u64, nullptr, 4, nullptr);
  }, "max_steps out of range");
  add("count: n -1", [=] {
    bh_trace_params p = good_params();
    return bh_count_outside_f64(&p, d, 4, -1, i32, nullptr);
  }, "bad n");
  add("count: n INT64_MAX", [=] {
    bh_trace_params p = good_params();
    return bh_count_outside_f64(&p, d, BIG, BIG, i32, nullptr);
  }, "bad n");
  add("count: ld < n", [=] {
    bh_trace_params p = good_params();
    return bh_count_outside_f64(&p, d, 3, 4, i32, nullptr);
  }, "bad n / ld");
  add("count: count NULL", [=] {
    bh_trace_params p = good_params();
    return bh_count_outside_f64(&p, d, 4, 4, nullptr, nullptr);
  }, "NULL buffer");

  // --- plain tracer ---
  auto plain = [=](const bh_tune_params *t, const double *in, int64_t ld, int64_t n) {
    bh_trace_params p = good_params();
    return bh_trace_plain_f64(&p, t, in, ld, d, ld, i32, i64, i64, u64, nullptr, n, nullptr);
  };
  add("plain: n 2^36+1", [=] { return plain(nullptr, d, BIG, ((int64_t)1 << 36) + 1); }, "bad n");
  add("plain: n INT64_MAX", [=] { return plain(nullptr, d, BIG, BIG); }, "bad n");
  add("plain: NULL input", [=] { return plain(nullptr, nullptr, 4, 4); }, "NULL buffer");
  add("plain: block 32", [=] {
    bh_tune_params t = tune_of(32, 1, 1);
    return plain(&t, d, 4, 4);
  }, "block_size must be");
  add("plain: block 320", [=] {
    bh_tune_params t = tune_of(320, 1, 1);
    return plain(&t, d, 4, 4);
  }, "block_size must be");
  add("plain: block -64", [=] {
    bh_tune_params t = tune_of(-64, 1, 1);
    return plain(&t, d, 4, 4);
  }, "block_size must be");
  add("plain: rays_per_lane -3", [=] {
    bh_tune_params t = tune_of(128, -3, 1);
    return plain(&t, d, 4, 4);
  }, "rays_per_lane");
  add("plain: refill 65", [=] {
    bh_tune_params t = tune_of(128, 1, 65);
    return plain(&t, d, 4, 4);
  }, "refill_min");
  add("plain: refill -1", [=] {
    bh_tune_params t = tune_of(128, 1, -1);
    return plain(&t, d, 4, 4);
  }, "refill_min");

  // --- batched tracer ---
  auto batched = [=](const bh_trace_params *jobs, int32_t n_jobs, int32_t mode,
                     const bh_tune_params *t, int64_t ld, int64_t n) {
    return bh_trace_batched_f64(jobs, i64, n_jobs, mode, t, d, ld, d, ld, i32, i64, i64, u64,
                                nullptr, u64, n, nullptr);
  };
  add("batched: 0 jobs", [=] {
    bh_trace_params p = good_params();
    return batched(&p, 0, BH_MODE_FAST, nullptr, 4, 4);
  }, "n_jobs");
  add("batched: 257 jobs", [=] {
    bh_trace_params p = good_params();
    return batched(&p, 257, BH_MODE_FAST, nullptr, 4, 4);
  }, "n_jobs");
  add("batched: mode 3", [=] {
    bh_trace_params p = good_params();
    return batched(&p, 1, 3, nullptr, 4, 4);
  }, "mode must be");
  add("batched: n_total INT64_MAX", [=] {
    bh_trace_params p = good_params();
    return batched(&p, 1, BH_MODE_FAST, nullptr, BIG, BIG);
  }, "bad n");
  add("batched: jobs NULL", [=] { return batched(nullptr, 1, BH_MODE_FAST, nullptr, 4, 4); },
      "NULL buffer");
  add("batched: block 96", [=] {
    bh_trace_params p = good_params();
    bh_tune_params t = tune_of(96, 0, 0);
    return batched(&p, 1, BH_MODE_FAST, &t, 4, 4);
  }, "block_size must be");
  add("batched: lanes_per_ray 2", [=] {
    bh_trace_params p = good_params();
    bh_tune_params t = tune_of(256, 0, 0);
    t.lanes_per_ray = 2;
    return batched(&p, 1, BH_MODE_PRECISE, &t, 4, 4);
  }, "lanes_per_ray must be 1 or 4");
  add("batched: quad lanes in the fast build", [=] {
    bh_trace_params p = good_params();
    bh_tune_params t = tune_of(256, 0, 0);
    t.lanes_per_ray = 4;
    return batched(&p, 1, BH_MODE_FAST, &t, 4, 4);
  }, "needs the precise build");
  add("batched: quad refill_min 17", [=] {
    bh_trace_params p = good_params();
    bh_tune_params t = tune_of(256, 0, 17);
    t.lanes_per_ray = 4;
    return batched(&p, 1, BH_MODE_PRECISE, &t, 4, 4);
  }, "refill_min must be in [1, 16]");
  add("plain: lanes_per_ray 4", [=] {
    bh_tune_params t = tune_of(128, 1, 1);
    t.lanes_per_ray = 4;
    return plain(&t, d, 4, 4);
  }, "lanes_per_ray 1");

  // --- adjoint record / backward ---
  auto record = [=](int32_t K, int32_t cap, double *ckpt) {
    bh_trace_params p = good_params();
    return bh_trace_record_f64(&p, d, 4, d, 4, i32, i64, i64, u64, nullptr, 4, K, cap, 20.0, ckpt,
                               i32, d, d, d, i64, d, i32, u64, nullptr);
  };
  add("record: K 0", [=] { return record(0, 4, d); }, "1 <= K <= 2^20");
  add("record: K 2^20+1", [=] { return record((1 << 20) + 1, 4, d); }, "1 <= K <= 2^20");
  add("record: cap 0", [=] { return record(8, 0, d); }, "cap >= 1");
  add("record: ckpt NULL", [=] { return record(8, 4, nullptr); }, "NULL buffer");
  auto adjoint = [=](int64_t n, int32_t K, double *scratch) {
    bh_trace_params p = good_params();
    return bh_trace_adjoint_f64(&p, n, K, 20.0, d, i32, i64, d, d, d, d, d, d, scratch, nullptr,
                                nullptr);
  };
  add("adjoint: n -5", [=] { return adjoint(-5, 8, d); }, "bad n");
  add("adjoint: n INT64_MAX", [=] { return adjoint(BIG, 8, d); }, "bad n");
  add("adjoint: K INT32_MAX", [=] { return adjoint(4, INT32_MAX, d); }, "1 <= K <= 2^20");
  add("adjoint: scratch NULL", [=] { return adjoint(4, 8, nullptr); }, "NULL buffer");
  auto record_traj = [=](int64_t n, double *pool, int64_t pages, int64_t *link, int64_t *last) {
    bh_trace_params p = good_params();
    return bh_trace_record_trajectory_f64(&p, d, 4, d, 4, i32, i64, i64, u64, nullptr, n, pool,
                                          pages, link, last, 20.0, d, d, d, i64, d, i32, u64,
                                          nullptr);
  };
  add("record trajectory: n -1", [=] { return record_traj(-1, d, 2, i64, i64); }, "bad n / ld");
  add("record trajectory: n > ld", [=] { return record_traj(5, d, 2, i64, i64); },
      "bad n / ld");
  add("record trajectory: pages -1", [=] { return record_traj(4, d, -1, i64, i64); },
      "pool_pages out of range");
  add("record trajectory: pages 2^48+1", [=] {
    return record_traj(4, d, ((int64_t)1 << 48) + 1, i64, i64);
  }, "pool_pages out of range");
  add("record trajectory: pool NULL", [=] { return record_traj(4, nullptr, 2, i64, i64); },
      "NULL buffer");
  add("record trajectory: page_link NULL", [=] { return record_traj(4, d, 2, nullptr, i64); },
      "NULL buffer");
  add("record trajectory: last_page NULL", [=] { return record_traj(4, d, 0, i64, nullptr); },
      "NULL buffer");
  auto adjoint_traj = [=](int64_t n, const double *pool, int64_t pages, const int64_t *last,
                          const int64_t *steps) {
    bh_trace_params p = good_params();
    return bh_trace_adjoint_trajectory_f64(&p, n, 20.0, pool, pages, i64, last, steps, i32, i64,
                                           d, d, d, d, d, d, nullptr, nullptr);
  };
  add("adjoint trajectory: n INT64_MAX", [=] { return adjoint_traj(BIG, d, 2, i64, i64); },
      "bad n");
  add("adjoint trajectory: pages -5", [=] { return adjoint_traj(4, d, -5, i64, i64); },
      "pool_pages out of range");
  add("adjoint trajectory: pool NULL", [=] { return adjoint_traj(4, nullptr, 2, i64, i64); },
      "NULL buffer");
  add("adjoint trajectory: last_page NULL", [=] { return adjoint_traj(4, d, 2, nullptr, i64); },
      "NULL buffer");
  add("adjoint trajectory: steps NULL", [=] { return adjoint_traj(4, d, 2, i64, nullptr); },
      "NULL buffer");
  add("adjoint trajectory: mode 3", [=] {
    bh_trace_params p = good_params();
    p.mode = 3;
    return bh_trace_adjoint_trajectory_f64(&p, 4, 20.0, d, 2, i64, i64, i64, i32, i64, d, d, d, d,
                                           d, d, nullptr, nullptr);
  }, "mode must be");
  auto radial = [=](double mass, int64_t n, int32_t grid, int32_t refine, const double *spin) {
    return bh_radial_minimum_f64(mass, spin, d, d, d, n, grid, refine, d, i32, nullptr);
  };
  add("radial_minimum: mass 0", [=] { return radial(0.0, 4, 256, 60, d); },
      "mass must be positive");
  add("radial_minimum: mass inf", [=] { return radial(__builtin_inf(), 4, 256, 60, d); },
      "mass must be positive");
  add("radial_minimum: n -1", [=] { return radial(1.0, -1, 256, 60, d); }, "bad n");
  add("radial_minimum: grid 1", [=] { return radial(1.0, 4, 1, 60, d); }, "2 <= grid");
  add("radial_minimum: refine -1", [=] { return radial(1.0, 4, 256, -1, d); }, "0 <= refine");
  add("radial_minimum: spin NULL", [=] { return radial(1.0, 4, 256, 60, nullptr); },
      "NULL buffer");
  add("rhs_vjp: spin NaN", [=] {
    return bh_rhs_vjp_f64(1.0, __builtin_nan(""), d, d, 4, d, d, nullptr);
  }, "Kerr bound");
  add("rhs_vjp: v NULL", [=] { return bh_rhs_vjp_f64(1.0, 0.5, d, nullptr, 4, d, d, nullptr); },
      "NULL buffer");

  // --- layout, geometry, single steps ---
  add("aos_to_soa: ld < n", [=] { return bh_aos_to_soa_f64(d, d, 2, 4, nullptr); }, "bad n / ld");
  add("aos_to_soa: NULL", [=] { return bh_aos_to_soa_f64(nullptr, d, 4, 4, nullptr); },
      "NULL buffer");
  add("soa_to_aos: n INT64_MAX", [=] { return bh_soa_to_aos_f64(d, BIG, d, BIG, nullptr); },
      "bad n");
  add("geometry: mode 2", [=] {
    return bh_geometry_f64(1.0, 0.5, 2, d, 4, d, d, d, d, d, nullptr);
  }, "mode must be");
  add("geometry: xyz NULL", [=] {
    return bh_geometry_f64(1.0, 0.5, BH_MODE_FAST, nullptr, 4, d, d, d, d, d, nullptr);
  }, "NULL buffer");
  add("geometry: n INT64_MAX", [=] {
    return bh_geometry_f64(1.0, 0.5, BH_MODE_FAST, d, BIG, d, d, d, d, d, nullptr);
  }, "bad n");
  add("rhs: mode -1", [=] { return bh_geodesic_rhs_f64(1.0, 0.5, -1, d, 4, d, nullptr); },
      "mode must be");
  add("rhs: deriv NULL", [=] {
    return bh_geodesic_rhs_f64(1.0, 0.5, BH_MODE_PRECISE, d, 4, nullptr, nullptr);
  }, "NULL buffer");
  add("dp5: mode 5", [=] {
    return bh_dp5_step_f64(1.0, 0.5, 5, 1e-8, 1e-10, d, d, 4, d, d, d, d, nullptr);
  }, "mode must be");
  add("dp5: h NULL", [=] {
    return bh_dp5_step_f64(1.0, 0.5, BH_MODE_FAST, 1e-8, 1e-10, d, nullptr, 4, d, d, d, d,
                           nullptr);
  }, "NULL buffer");

  // --- kernel attributes ---
  add("attributes: out NULL", [=] { return bh_kernel_attributes(0, 0, 64, nullptr); },
      "out is NULL");
  add("attributes: mode 9", [=] { return bh_kernel_attributes(0, 9, 64, i32); }, "bad mode");
  add("attributes: kernel 7", [=] { return bh_kernel_attributes(7, 0, 64, i32); }, "bad kernel");
  add("attributes: quad kernel, fast build", [=] {
    return bh_kernel_attributes(BH_KERNEL_QUAD, BH_MODE_FAST, 64, i32);
  }, "precise build only");
  add("attributes: kernel -1", [=] { return bh_kernel_attributes(-1, 0, 64, i32); },
      "bad kernel");
  add("attributes: block 0", [=] { return bh_kernel_attributes(0, 0, 0, i32); }, "bad block");
  add("attributes: block 320", [=] { return bh_kernel_attributes(0, 0, 320, i32); },
      "bad block");
  add("attributes: adjoint block 128", [=] {
    return bh_kernel_attributes(BH_KERNEL_ADJOINT, 0, 128, i32);
  }, "64-thread blocks");

  // --- shader-like fixed-step tracer ---
  auto shader = [=](const bh_shader_params *p, int64_t ld_out, double *out, int32_t *status) {
    return bh_trace_shader_f64(p, d, 4, out, ld_out, status, d, nullptr, 4, nullptr);
  };
  add("shader: params NULL", [=] { return shader(nullptr, 4, d, i32); }, "params is NULL");
  add("shader: spin 2", [=] {
    bh_shader_params s = good_shader();
    s.spin = 2.0;
    return shader(&s, 4, d, i32);
  }, "Kerr bound");
  add("shader: mode 3", [=] {
    bh_shader_params s = good_shader();
    s.mode = 3;
    return shader(&s, 4, d, i32);
  }, "bad mode");
  add("shader: max_steps -1", [=] {
    bh_shader_params s = good_shader();
    s.max_steps = -1;
    return shader(&s, 4, d, i32);
  }, "bad max_steps");
  add("shader: max_steps INT64_MAX", [=] {
    bh_shader_params s = good_shader();
    s.max_steps = BIG;
    return shader(&s, 4, d, i32);
  }, "bad max_steps");
  add("shader: flags 4", [=] {
    bh_shader_params s = good_shader();
    s.flags = 4;
    return shader(&s, 4, d, i32);
  }, "bad flags");
  add("shader: ld_out < n", [=] {
    bh_shader_params s = good_shader();
    return shader(&s, 2, d, i32);
  }, "bad n / ld");
  add("shader: status NULL", [=] {
    bh_shader_params s = good_shader();
    return shader(&s, 4, d, nullptr);
  }, "NULL buffer");

  // --- fixed-step integrators ---
  add("midpoint: iterations -1", [=] {
    return bh_midpoint_f64(1.0, 0.5, BH_MODE_FAST, d, 4, d, 4, 10, -1, nullptr);
  }, "bad n / ld / n_steps / iterations");
  add("midpoint: n_steps -1", [=] {
    return bh_midpoint_f64(1.0, 0.5, BH_MODE_FAST, d, 4, d, 4, -1, 4, nullptr);
  }, "bad n / ld / n_steps / iterations");
  add("midpoint: mode 2", [=] {
    return bh_midpoint_f64(1.0, 0.5, 2, d, 4, d, 4, 10, 4, nullptr);
  }, "bad mode");
  add("rk4: n_steps -1", [=] {
    return bh_rk4_f64(1.0, 0.5, BH_MODE_FAST, d, 4, d, 4, -1, nullptr);
  }, "bad n / ld / n_steps");
  add("rk4: h NULL", [=] {
    return bh_rk4_f64(1.0, 0.5, BH_MODE_FAST, d, 4, nullptr, 4, 10, nullptr);
  }, "NULL buffer");
  add("rk4: n INT64_MAX", [=] {
    return bh_rk4_f64(1.0, 0.5, BH_MODE_FAST, d, BIG, d, BIG, 10, nullptr);
  }, "bad n / ld / n_steps");
  add("rk4: mode 4", [=] { return bh_rk4_f64(1.0, 0.5, 4, d, 4, d, 4, 10, nullptr); },
      "bad mode");

  // --- offline tile tracer ---
  auto tile = [=](const bh_tile_params *t, const bh_tune_params *tune, uint8_t *sat) {
    bh_trace_params p = good_params();
    return bh_trace_tile_f64(&p, t, tune, d, 4, d, 4, i32, i64, i64, sat, d, d, d, u64, nullptr,
                             4, nullptr);
  };
  bh_tile_params tp;
  memset(&tp, 0, sizeof(tp));
  tp.disk_inner = 6.0;
  tp.disk_outer = 20.0;
  tp.saturate_steps = 850;
  tp.flags = BH_TILE_DISK;
  add("tile: params NULL", [=] { return tile(nullptr, nullptr, g_b); }, "NULL tile parameters");
  add("tile: flags 8", [=] {
    bh_tile_params t = tp;
    t.flags = 8;
    return tile(&t, nullptr, g_b);
  }, "bad tile flags");
  add("tile: disk inner > outer", [=] {
    bh_tile_params t = tp;
    t.disk_inner = 30.0;
    return tile(&t, nullptr, g_b);
  }, "disk_inner must not exceed");
  add("tile: disk inner NaN", [=] {
    bh_tile_params t = tp;
    t.disk_inner = __builtin_nan("");
    return tile(&t, nullptr, g_b);
  }, "disk_inner must not exceed");
  add("tile: block 100", [=] {
    bh_tune_params t = tune_of(100, 0, 0);
    return tile(&tp, &t, g_b);
  }, "block_size must be");
  add("tile: saturated NULL", [=] { return tile(&tp, nullptr, nullptr); }, "NULL buffer");

  // --- dense-output integrator ---
  auto dense = [=](int64_t n, int64_t cap, double rtol, double atol, double max_step,
                   double lambda_total, double *ctl) {
    return bh_integrate_dense_f64(1.0, 0.5, BH_MODE_PRECISE, rtol, atol, max_step, lambda_total,
                                  d, n, n, d, d, cap, ctl, i64, i32, nullptr);
  };
  add("dense: n 0", [=] { return dense(0, 8, 1e-9, 1e-12, 1.0, 10.0, d); }, "need n >= 1");
  add("dense: n INT64_MAX", [=] { return dense(BIG, 8, 1e-9, 1e-12, 1.0, 10.0, d); },
      "need n >= 1");
  add("dense: cap 0", [=] { return dense(2, 0, 1e-9, 1e-12, 1.0, 10.0, d); }, "cap >= 1");
  add("dense: rtol -1", [=] { return dense(2, 8, -1.0, 1e-12, 1.0, 10.0, d); },
      "bad tolerances");
  add("dense: atol NaN", [=] { return dense(2, 8, 1e-9, __builtin_nan(""), 1.0, 10.0, d); },
      "bad tolerances");
  add("dense: max_step 0", [=] { return dense(2, 8, 1e-9, 1e-12, 0.0, 10.0, d); },
      "bad tolerances");
  add("dense: lambda NaN", [=] { return dense(2, 8, 1e-9, 1e-12, 1.0, __builtin_nan(""), d); },
      "bad tolerances");
  add("dense: ctl NULL", [=] { return dense(2, 8, 1e-9, 1e-12, 1.0, 10.0, nullptr); },
      "NULL buffer");

  // --- camera rays and shading ---
  add("camera: camera NULL", [=] {
    return bh_camera_rays_f64(1.0, 0.5, nullptr, 0.2, 0.1, 4, 4, 1, d, 16, nullptr);
  }, "camera is NULL");
  add("camera: width -1", [=] {
    return bh_camera_rays_f64(1.0, 0.5, d, 0.2, 0.1, -1, 4, 1, d, 16, nullptr);
  }, "bad width / height");
  add("camera: width * height overflows", [=] {
    return bh_camera_rays_f64(1.0, 0.5, d, 0.2, 0.1, BIG, 3, 1, d, BIG, nullptr);
  }, "bad width / height");
  add("camera: 2^19 x 2^18", [=] {
    return bh_camera_rays_f64(1.0, 0.5, d, 0.2, 0.1, (int64_t)1 << 19, (int64_t)1 << 18, 1, d,
                              BIG, nullptr);
  }, "bad width / height");
  add("camera: ld < width * height", [=] {
    return bh_camera_rays_f64(1.0, 0.5, d, 0.2, 0.1, 4, 4, 1, d, 15, nullptr);
  }, "ld < width * height");
  add("shade: cells 0", [=] { return bh_shade_checker_u8(1.0, 0.5, d, 4, i32, 4, 0, g_b, nullptr); },
      "bad n / ld / cells");
  add("shade: rgb NULL", [=] {
    return bh_shade_checker_u8(1.0, 0.5, d, 4, i32, 4, 8, nullptr, nullptr);
  }, "NULL buffer");
  add("checker colours: cells -3", [=] { return bh_checker_colors_u8(d, 4, -3, g_b, nullptr); },
      "bad n / cells");
  add("checker colours: n INT64_MAX", [=] {
    return bh_checker_colors_u8(d, BIG, 12, g_b, nullptr);
  }, "bad n / cells");
  add("checker colours: directions NULL", [=] {
    return bh_checker_colors_u8(nullptr, 4, 12, g_b, nullptr);
  }, "NULL buffer");

  // --- controller helpers ---
  add("error ratio: mode 9", [=] {
    return bh_error_ratio_f64(9, d, d, d, 4, 8, 1e-8, 1e-10, d, nullptr);
  }, "mode must be");
  add("error ratio: n INT64_MAX", [=] {
    return bh_error_ratio_f64(BH_MODE_PRECISE, d, d, d, BIG, 8, 1e-8, 1e-10, d, nullptr);
  }, "bad n");
  add("error ratio: dim 129", [=] {
    return bh_error_ratio_f64(BH_MODE_FAST, d, d, d, 4, 129, 1e-8, 1e-10, d, nullptr);
  }, "dim must be");
  add("error ratio: ratio NULL", [=] {
    return bh_error_ratio_f64(BH_MODE_PRECISE, d, d, d, 4, 2, 1e-8, 1e-10, nullptr, nullptr);
  }, "NULL buffer");
  add("step factor: mode -1", [=] { return bh_step_factor_f64(-1, d, 4, d, nullptr); },
      "mode must be");
  add("step factor: n -4", [=] { return bh_step_factor_f64(BH_MODE_FAST, d, -4, d, nullptr); },
      "bad n");
  add("step factor: factor NULL", [=] {
    return bh_step_factor_f64(BH_MODE_PRECISE, d, 4, nullptr, nullptr);
  }, "NULL buffer");

  // --- momenta and diagnostics ---
  add("null momentum: n -1", [=] {
    return bh_null_momentum_f64(1.0, 0.5, d, d, -1, 1, 1, d, nullptr);
  }, "bad n");
  add("null momentum: momenta NULL", [=] {
    return bh_null_momentum_f64(1.0, 0.5, d, d, 4, 1, 1, nullptr, nullptr);
  }, "NULL buffer");
  add("timelike: velocities NULL", [=] {
    return bh_timelike_momentum_f64(1.0, 0.5, d, nullptr, 4, d, i32, nullptr);
  }, "NULL buffer");
  add("timelike: mass -1", [=] {
    return bh_timelike_momentum_f64(-1.0, 0.5, d, d, 4, d, i32, nullptr);
  }, "mass must be positive");
  add("diagnostics: state NULL", [=] {
    return bh_diagnostics_f64(1.0, 0.5, nullptr, 4, d, d, d, d, nullptr);
  }, "NULL buffer");
  add("diagnostics: n INT64_MAX", [=] {
    return bh_diagnostics_f64(1.0, 0.5, d, BIG, d, d, d, d, nullptr);
  }, "bad n");
  return v;
}

int check_case(const Case &c, bool verbose) {
  const int32_t rc = c.call();
  const char *msg = bh_last_error();
  if (rc != BH_EINVAL || !msg || !strstr(msg, c.message)) {
    fprintf(stderr, "FAIL %s: rc=%d message=\"%s\" (expected BH_EINVAL and \"%s\")\n",
            c.name.c_str(), (int)rc, msg ? msg : "(null)", c.message);
    return 1;
  }
  if (verbose) printf("ok   %s: %s\n", c.name.c_str(), msg);
  return 0;
}

int run_cases() {
  int failures = 0;
  if (bh_abi_version() != BH_ABI_VERSION) {
    fprintf(stderr, "FAIL abi version %d\n", (int)bh_abi_version());
    ++failures;
  }
  const struct {
    int32_t K, expect;
  } scratch[] = {{1, 9}, {64, 576}, {1 << 20, 9 << 20}, {0, -1}, {-7, -1},
                 {(1 << 20) + 1, -1}, {INT32_MAX, -1}, {INT32_MIN, -1}};
  for (const auto &s : scratch) {
    if (bh_adjoint_scratch_doubles(s.K) != s.expect) {
      fprintf(stderr, "FAIL bh_adjoint_scratch_doubles(%d) = %d, expected %d\n", (int)s.K,
              (int)bh_adjoint_scratch_doubles(s.K), (int)s.expect);
      ++failures;
    }
  }
  const std::vector<Case> cases = build_cases();
  for (const Case &c : cases) failures += check_case(c, true);
  printf("%zu cases, %d failures\n", cases.size(), failures);
  return failures ? 1 : 0;
}

int run_threads() {
  const std::vector<Case> cases = build_cases();
  const int n_threads = 8, rounds = 400;
  std::atomic<int> failures{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < n_threads; ++t) {
    pool.emplace_back([&, t] {
      // each thread walks the case list from its own offset, so at any time
      // the threads are refusing different calls with different messages
      for (int r = 0; r < rounds; ++r) {
        const Case &c = cases[(size_t)(t * 11 + r) % cases.size()];
        const int32_t rc = c.call();
        const std::string mine = bh_last_error();
        std::this_thread::yield();
        if (rc != BH_EINVAL || mine.find(c.message) == std::string::npos ||
            mine != bh_last_error()) {
          fprintf(stderr, "FAIL thread %d round %d %s: \"%s\"\n", t, r, c.name.c_str(),
                  mine.c_str());
          failures.fetch_add(1);
          return;
        }
      }
    });
  }
  for (auto &th : pool) th.join();
  printf("%d threads x %d refusals, %d failures\n", n_threads, rounds, failures.load());
  return failures.load() ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "cases";
  if (mode == "cases") return run_cases();
  if (mode == "threads") return run_threads();
  fprintf(stderr, "usage: %s [cases|threads]\n", argv[0]);
  return 2;
}
