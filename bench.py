"""Benchmark: FP64 adaptive DP5(4) ray tracing on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hd|small]

One bench step traces the whole 1920x1080 sweep -- spins a in {0, 0.5, 0.9,
0.998} x inclinations {17, 80} deg, camera at r = 50 M, horizontal FOV 25 deg,
rtol 1e-8 / atol 1e-10, capture at r_+ + 0.01 M, escape at 100 M, max_steps
20000 -- one plain-tracer launch per frame (8 launches, 16.6 M rays), rays
handed to wavefront lanes in 8x8 pixel tiles, the tiles nearest the image
centre (shadow and photon ring, the longest traces) first (--order; results
do not depend on it).  Rays are generated on the device before timing (camera directions
on the host, null momenta and the SoA layout on the device), so the timed
region is the tracer kernels only, with inputs resident in HBM.

Prints one JSON line.  value = accepted ray-steps per second over the sweep.
roofline: algorithmic FP64 FLOP of the tracer kernel (1321 FLOP per DP5
attempt: 6 RHS x 146 + 406 stage + 36 error-norm + 3 controller) per launch,
over the kernel's mean launch time measured with HIP events on its stream,
against the FP64 vector peak (78.6 TFLOP/s).  cpu_baseline: the NumPy oracle
(a restatement of the reference tracer) on a seeded subsample of every frame,
on this host's cores, timed before the GPU is touched.  adjoint_over_forward:
the discrete-adjoint op's time over the plain forward trace's on the 512^2
shadow fit (--workload shadowfit), measured after the sweep.  Also after the
sweep (``targets``; --no-targets skips them): batched_over_per_job, the
uneven 64-job mix as one persistent batched plan against one plain launch
per job, both autotuned (--workload uneven); the 65,536-orbit timelike
workload's forward time, E / L_z / H drift and adjoint/forward (--workload
orbits); the 512^2 spin and inclination fit (--workload fit); and one a = 0.9
frame's ray generation (host directions, device states) and host<->device
copies, reported apart from the kernel time.
Launch times are reported as means (roofline, matching rocprof's average)
and medians.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FLOP_PER_ATTEMPT = 1321
FP64_VECTOR_PEAK_TFLOPS = 78.6
# HBM bytes per trace_plain_kernel launch on the 1920x1080 frames (fast
# build): rocprofv3 FETCH_SIZE + WRITE_SIZE, each calibrated on the layout
# copy aos_to_soa_kernel of the same run (profiles/trace_plain_hd_sweep.md).
HD_TRAFFIC_BYTES_PER_LAUNCH = 3.982e8
METRIC = "ray-steps/s FP64 @rtol 1e-8, % FP64 vector peak, adjoint/forward time"


def frames(workload: str):
    from blackhorizon_amd import workloads

    if workload == "hd":
        return workloads.hd_frames()
    if workload == "hd09":
        return [workloads.hd_frame(0.9, 17.0), workloads.hd_frame(0.9, 80.0)]
    if workload == "small":
        return [workloads.small_frame()]
    raise SystemExit(f"unknown workload {workload}")


def camera_arrays(job):
    from blackhorizon_amd.camera import PinholeCamera

    cam = PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                   azimuth_degrees=job.azimuth_degrees,
                                   fov_degrees=job.fov_degrees, width=job.width,
                                   height=job.height)
    return cam.ray_origins(), cam.ray_directions()


# --------------------------------------------------------------------------
# CPU baseline (NumPy oracle), run before any GPU work
# --------------------------------------------------------------------------

def _cpu_shard(args):
    job_dict, pos, dirs = args
    from blackhorizon_amd.workloads import CameraJob
    from oracle import bh_oracle as orc

    job = CameraJob(**job_dict)
    st = orc.Kerr(job.mass, job.spin)
    s0 = orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, dirs, "past"))
    t0 = time.perf_counter()
    res, att = orc.trace_rays(st, s0, job.escape_radius, max_steps=job.max_steps,
                              rtol=job.rtol, atol=job.atol, initial_step=job.initial_step,
                              max_step=job.max_step, min_step=job.min_step,
                              capture_margin=job.capture_margin, return_attempts=True)
    return time.perf_counter() - t0, int(res.steps.sum()), int(att.sum())


def cpu_baseline(jobs, per_frame: int, seed: int = 7, processes: int = 1):
    shards = []
    for f, job in enumerate(jobs):
        pos, dirs = camera_arrays(job)
        rng = np.random.default_rng(seed + f)
        idx = np.sort(rng.choice(job.n_rays, size=min(per_frame, job.n_rays), replace=False))
        for p in range(processes):
            sl = idx[p::processes]
            shards.append((job.as_dict(), pos[sl], dirs[sl]))
    t0 = time.perf_counter()
    if processes == 1:
        out = [_cpu_shard(s) for s in shards]
    else:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(processes) as pool:
            out = pool.map(_cpu_shard, shards)
    wall = time.perf_counter() - t0
    steps = sum(o[1] for o in out)
    attempts = sum(o[2] for o in out)
    return dict(value=steps / wall, unit="ray-steps/s", cores=processes, kind="port",
                sample=f"{min(per_frame, jobs[0].n_rays)} seeded rays of each of {len(jobs)} "
                       f"frames ({len(jobs) * min(per_frame, jobs[0].n_rays)} rays), NumPy "
                       f"oracle trace_rays, {processes} process(es)",
                wall_s=round(wall, 3), ray_steps=steps, attempts=attempts,
                ns_per_ray_step=wall / max(steps, 1) * 1e9)


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------

def camera_order(job, order: str):
    if order == "row":
        return None
    from blackhorizon_amd.camera import PinholeCamera

    cam = PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                   azimuth_degrees=job.azimuth_degrees,
                                   fov_degrees=job.fov_degrees, width=job.width,
                                   height=job.height)
    if order.startswith("center"):
        return cam.center_tile_order(int(order[6:]) if order[6:] else 8)
    return cam.tile_order(int(order[4:]) if order.startswith("tile") and order[4:] else 8)


def device_state(job):
    """(n, 8) camera rays of a job built on the device."""
    import torch

    from blackhorizon_amd import geodesics, kerr

    dev = torch.device("cuda", 0)
    st = kerr.KerrSpacetime(job.mass, job.spin)
    pos, dirs = camera_arrays(job)
    pos_t = torch.from_numpy(pos).to(dev)
    mom = geodesics.null_momentum_from_velocity(st, pos_t, torch.from_numpy(dirs).to(dev), "past")
    return st, geodesics.build_state(pos_t, mom)


def prepare_gpu(jobs, mode, order="row"):
    """One device-resident TracePlan per job (rays built on the device)."""
    import torch

    from blackhorizon_amd import tracer

    prepared = []
    for job in jobs:
        st, state = device_state(job)
        o = camera_order(job, order)
        plan = tracer.TracePlan(st, state, job.escape_radius, job.max_steps, job.rtol, job.atol,
                                job.initial_step, job.max_step, job.min_step,
                                job.capture_margin, mode=mode,
                                ray_order=None if o is None else torch.from_numpy(o))
        plan.state0 = None  # the SoA copy is all the kernel reads
        del state
        prepared.append((job, plan))
    torch.cuda.synchronize()
    return prepared


def lane_efficiency(plan, rays_per_lane):
    """Fraction of lane-attempts doing useful work when every 64 consecutive
    rays of the hand-out order share a wavefront (exact for rays_per_lane 1)."""
    import torch

    a = plan.attempts if plan.order is None else plan.attempts[plan.order]
    n = a.shape[0]
    pad = (-n) % 64
    if pad:
        a = torch.cat([a, torch.zeros(pad, dtype=a.dtype, device=a.device)])
    w = a.view(-1, 64)
    return float(a.sum().item()) / float((64 * w.max(dim=1).values).sum().item())


def pick_tune(args, prepared, kind_key, workload=None):
    """--tune: "auto" runs the autotuner on the first plan (its table goes
    into the JSON line), then times its three fastest shapes on every plan
    of the workload (``workload``, default ``prepared``) and keeps the one
    with the least total (shapes within noise on one frame can differ by
    several per cent over the sweep);
    "default" is (256, 1, 1), "B,R,M" is explicit."""
    from blackhorizon_amd import autotune, tracer

    if args.tune == "default":
        return tracer.Tune(), None
    if args.tune != "auto":
        b, r, m = (int(v) for v in args.tune.split(","))
        return tracer.Tune(b, r, m), None
    plan = prepared[0][1]
    res = autotune.autotune(plan, workload_key=kind_key, cache=None, reps=2)
    print(f"[bench] autotune: {res.best} ({res.best_ms:.3f} ms)", file=sys.stderr, flush=True)
    ranked = sorted(res.table, key=lambda e: e[1])
    best = res.best
    totals = []
    workload = workload or prepared
    if len(workload) > 1:
        for t, _ in ranked[:3]:
            tune = tracer.Tune(**t)
            total = sum(autotune.time_candidates(p, [tune], reps=2)[0][1] for _, p in workload)
            totals.append([t, round(total, 4)])
        best = tracer.Tune(**min(totals, key=lambda e: e[1])[0])
        print(f"[bench] over the workload: {best}", file=sys.stderr, flush=True)
    info = dict(autotuned_on=kind_key, best_ms=res.best_ms,
                candidates=len(res.table),
                fastest5=[[t, round(ms, 4)] for t, ms in ranked[:5]],
                workload_ms_of_fastest3=totals)
    return best, info


def frame_io(job, mode: str, tune, reps: int = 5) -> dict:
    """One 1920x1080 frame's costs outside the tracer kernel, each the median
    of ``reps`` runs: camera directions on the host (NumPy, as the reference
    builds them, camera.py:96-113), the whole state built on the device
    (camera_rays_kernel), the (n, 8) initial states copied host -> device
    from pageable and from pinned memory, and the results (states, status,
    steps) copied device -> host as the library returns them to NumPy
    callers (pinned host memory) and into pageable memory; plus the frame's
    trace kernel time and the ray-step rate with the library's copies
    included."""
    import torch

    from blackhorizon_amd import _dev, kerr, tracer
    from blackhorizon_amd.camera import PinholeCamera

    st = kerr.KerrSpacetime(job.mass, job.spin)
    cam = PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                   azimuth_degrees=job.azimuth_degrees,
                                   fov_degrees=job.fov_degrees, width=job.width,
                                   height=job.height)
    cur = torch.cuda.current_stream()

    def dev_ms(fn):
        out = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(cur)
            fn()
            e1.record(cur)
            torch.cuda.synchronize()
            out.append(e0.elapsed_time(e1))
        return float(np.median(out))

    def host_ms(fn):
        out = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            out.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(out))

    ms_dirs_host = host_ms(cam.ray_directions)
    holder = {}
    ms_states_dev = dev_ms(lambda: holder.__setitem__("s", cam.device_states(st)))
    s_dev = holder["s"]
    s_host = s_dev.cpu().numpy()
    pinned = torch.from_numpy(s_host).pin_memory()
    ms_h2d = dev_ms(lambda: torch.from_numpy(s_host).to("cuda"))
    ms_h2d_pinned = dev_ms(lambda: pinned.to("cuda", non_blocking=True))
    plan = tracer.TracePlan(st, s_dev, job.escape_radius, job.max_steps, job.rtol, job.atol,
                            job.initial_step, job.max_step, job.min_step, job.capture_margin,
                            mode=mode, ray_order=torch.from_numpy(cam.center_tile_order(8)))
    plan.launch(tune)
    ms_kernel = dev_ms(lambda: plan.launch(tune))
    res = plan.result()  # device tensors, like the state it traced

    def d2h():  # the library's path for NumPy callers (pinned host memory)
        for t in (res.states, res.status, res.steps):
            _dev.to_numpy(t)

    def d2h_pageable():
        for t in (res.states, res.status, res.steps):
            t.cpu()
    ms_d2h = dev_ms(d2h)
    ms_d2h_pageable = dev_ms(d2h_pageable)
    steps = int(plan.steps.sum().item())
    # the whole drop-in call of a NumPy caller: upload, initial-radius check,
    # trace, results back as NumPy
    order = torch.from_numpy(cam.center_tile_order(8))
    ms_numpy_call = host_ms(lambda: tracer.trace_rays(
        st, s_host, job.escape_radius, job.max_steps, job.rtol, job.atol, job.initial_step,
        job.max_step, job.min_step, job.capture_margin, mode=mode, tune=tune, ray_order=order))
    return dict(spin=job.spin, inclination=job.inclination_degrees, rays=job.n_rays,
                ms_camera_directions_host=ms_dirs_host, ms_states_on_device=ms_states_dev,
                ms_h2d_states_pageable=ms_h2d, ms_h2d_states_pinned=ms_h2d_pinned,
                ms_d2h_results=ms_d2h, ms_d2h_results_pageable=ms_d2h_pageable,
                ms_trace_kernel=ms_kernel,
                ray_steps=steps, ray_steps_per_s_kernel=steps / (ms_kernel * 1e-3),
                ray_steps_per_s_with_copies=steps / ((ms_h2d + ms_kernel + ms_d2h) * 1e-3),
                ms_trace_rays_numpy=ms_numpy_call,
                ray_steps_per_s_trace_rays_numpy=steps / (ms_numpy_call * 1e-3))


def run_uneven(args) -> dict:
    """The uneven 64-job mix: one persistent batched launch per arithmetic
    mode (fast jobs and precise jobs on two streams) against one plain
    launch per job on one stream.  Same rays, same per-ray results."""
    import torch

    from blackhorizon_amd import _lib, autotune, batch, graph, tracer, workloads

    _lib.require_device()
    jobs = workloads.uneven_mix()
    tjobs = []
    for job in jobs:
        st, state = device_state(job)
        tjobs.append(batch.TraceJob(st, state, job.escape_radius, job.max_steps, job.rtol,
                                    job.atol, job.initial_step, job.max_step, job.min_step,
                                    job.capture_margin))
    plan = batch.BatchPlan(tjobs, mode=args.mode, concurrent=not args.serial_groups,
                           order_fn=batch.longest_first if args.batch_order == "longest_first"
                           else None)
    prepared = prepare_gpu(jobs, args.mode, "row")
    for t in tjobs:
        t.state0 = None
    torch.cuda.synchronize()
    cur = torch.cuda.current_stream()
    tune_info = None
    if args.tune == "auto":
        res = autotune.autotune(plan, workload_key="uneven64", cache=None,
                                reps=getattr(args, "tune_reps", 2))
        tune_info = dict(best=res.best.__dict__, best_ms=res.best_ms, candidates=len(res.table))
        print(f"[bench] autotune batched: {res.best}", file=sys.stderr, flush=True)
    elif args.tune != "default":
        b_, _, m_ = (int(v) for v in args.tune.split(","))
        plan.tune = tracer.Tune(b_, 1, m_)

    def timed(fn, k):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(cur)
        for _ in range(k):
            fn()
        e1.record(cur)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    class PerJob:
        """One plain launch per job, in job order on one stream."""

        tune = None  # the plain kernel's launch shape for every job

        def launch(self, tune=None):
            for _, p in prepared:
                p.launch(tune or self.tune)

    per_job = PerJob()
    per_job_tune_info = None
    if args.tune == "auto":
        # the baseline gets its own best launch shape too
        table = autotune.time_candidates(per_job, autotune.plain_candidates(),
                                         reps=getattr(args, "tune_reps", 2))
        best, best_ms = autotune.select(table)
        per_job.tune = best
        per_job_tune_info = dict(best=best.__dict__, best_ms=best_ms, candidates=len(table))
        print(f"[bench] autotune per-job plain: {best}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        plan.launch()
        per_job.launch()
    torch.cuda.synchronize()
    ms_batched = timed(plan.launch, args.steps)
    print("[bench] batched done", file=sys.stderr, flush=True)
    ms_per_job = timed(per_job.launch, args.steps)
    print("[bench] per-job done", file=sys.stderr, flush=True)
    # the same launch sequences captured in HIP graphs
    g_batched = graph.TraceGraph(plan)
    g_per_job = graph.TraceGraph(per_job)
    ms_batched_graph = timed(g_batched.replay, args.steps)
    ms_per_job_graph = timed(g_per_job.replay, args.steps)
    print("[bench] graph replays done", file=sys.stderr, flush=True)
    steps_b = sum(int(g.steps.sum().item()) for g in plan.groups)
    att_b = sum(int(g.attempts.sum().item()) for g in plan.groups)
    steps_p = sum(int(p.steps.sum().item()) for _, p in prepared)
    same = all(torch.equal(g.steps, torch.cat([prepared[i][1].steps for i in g.idx]))
               for g in plan.groups)
    modes = {g.mode: len(g.idx) for g in plan.groups}
    return dict(
        batched_ray_steps_per_s=steps_b / (ms_batched * 1e-3),
        per_job_ray_steps_per_s=steps_p / (ms_per_job * 1e-3),
        speedup=ms_per_job / ms_batched,
        ms_batched=ms_batched, ms_per_job=ms_per_job,
        ms_batched_graph=ms_batched_graph, ms_per_job_graph=ms_per_job_graph,
        speedup_graph=ms_per_job_graph / ms_batched_graph,
        tune=tune_info or plan.tune.__dict__,
        tune_per_job=per_job_tune_info or tracer.Tune().__dict__,
        ray_steps=steps_b, attempts=att_b, rays=plan.total_rays,
        pct_fp64_vector_peak_batched=100.0 * att_b * FLOP_PER_ATTEMPT / (ms_batched * 1e-3)
        / (FP64_VECTOR_PEAK_TFLOPS * 1e12),
        jobs_per_mode=modes, queue_order=args.batch_order,
        per_ray_steps_identical=bool(same and steps_b == steps_p))


RK4_FLOP_PER_STEP = 4 * 146 + 104  # 4 RHS + stage and update arithmetic, 8 components


def rk4_rays(n: int):
    """Positions and directions of the reference benchmark's rays
    (examples/benchmark.py:22-31): seed-7 isotropic unit vectors u, photons
    at 50 u aimed at -u."""
    rng = np.random.default_rng(7)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    return 50.0 * u, -u


def rk4_states_device(n: int):
    """(n, 8) states of the benchmark's rays built on the device (a = 0.9)."""
    import torch

    from blackhorizon_amd import geodesics, kerr

    pos, dirs = rk4_rays(n)
    st = kerr.KerrSpacetime(1.0, 0.9)
    p = torch.from_numpy(pos).cuda()
    return geodesics.build_state(p, geodesics.null_momentum_from_velocity(
        st, p, torch.from_numpy(dirs).cuda()))


def rk4_states_oracle(n: int):
    """The same rays built by the oracle (CPU baseline and checker)."""
    from oracle import bh_oracle as orc

    pos, dirs = rk4_rays(n)
    st = orc.Kerr(spin=0.9)
    return orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, dirs))


def run_rk4(args) -> dict:
    """The reference's own throughput benchmark, like for like: fixed-step
    RK4, h = 0.05, 50 steps of 1e5 rays (examples/benchmark.py:34-62), plus
    the same at 2^21 rays, which fills the GPU; all 50 steps run in one
    launch with the state in registers.  The reference publishes 1376-1700
    ns/ray-step for NumPy FP64 on a CPU and 242.2 ns for CuPy FP64 on an RTX
    3070 (docs/DESIGN.md:282-285) on this benchmark."""
    import torch

    from blackhorizon_amd import _lib, integrators, kerr

    lib = _lib.require_device()
    st = kerr.KerrSpacetime(1.0, 0.9)
    out = {}
    cpu = None
    for n in (100_000, 1 << 21):
        state = rk4_states_device(n)
        soa = torch.empty((8, n), dtype=torch.float64, device="cuda")
        stream = _lib.stream_handle()
        _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(state), _lib.ptr(soa), n, n, stream), "soa")
        h = torch.full((n,), 0.05, dtype=torch.float64, device="cuda")
        work = soa.clone()
        cur = torch.cuda.current_stream()
        for _ in range(args.warmup):
            work.copy_(soa)
            _lib.check(lib.bh_rk4_f64(1.0, 0.9, _lib.MODE_FAST, _lib.ptr(work), n, _lib.ptr(h), n,
                                      50, stream), "rk4")
        ms = []
        for _ in range(args.steps):
            work.copy_(soa)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(cur)
            _lib.check(lib.bh_rk4_f64(1.0, 0.9, _lib.MODE_FAST, _lib.ptr(work), n, _lib.ptr(h), n,
                                      50, stream), "rk4")
            e1.record(cur)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        t = float(np.median(ms))
        ray_steps = n * 50
        out[f"rays_{n}"] = dict(ms=t, ns_per_ray_step=t * 1e6 / ray_steps,
                                ray_steps_per_s=ray_steps / (t * 1e-3),
                                pct_fp64_vector_peak=100.0 * ray_steps * RK4_FLOP_PER_STEP
                                / (t * 1e-3) / (FP64_VECTOR_PEAK_TFLOPS * 1e12))
        # the implicit midpoint rule (integrators.py:177-207) on the same rays:
        # 50 steps of 4 fixed-point iterations each, fused in one launch
        mms = []
        for k in range(args.warmup + args.steps):
            work.copy_(soa)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(cur)
            _lib.check(lib.bh_midpoint_f64(1.0, 0.9, _lib.MODE_FAST, _lib.ptr(work), n,
                                           _lib.ptr(h), n, 50, 4, stream), "midpoint")
            e1.record(cur)
            e1.synchronize()
            if k >= args.warmup:
                mms.append(e0.elapsed_time(e1))
        tm = float(np.median(mms))
        out[f"midpoint_rays_{n}"] = dict(ms=tm, ns_per_ray_step=tm * 1e6 / ray_steps,
                                         ray_steps_per_s=ray_steps / (tm * 1e-3))
        print(f"[bench] rk4 {n} rays done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        from oracle import bh_oracle as orc

        # the precise build equals the NumPy restatement bit for bit (check)
        sample = rk4_states_oracle(4096)
        got = integrators.rk4_steps(integrators.GeodesicRHS(st, "precise"), sample,
                                    np.full(4096, 0.05), 50)
        ost = orc.Kerr(spin=0.9)
        y = rk4_states_oracle(20_000)
        hh = np.full(y.shape[0], 0.05)
        t0 = time.perf_counter()
        for _ in range(50):
            y = orc.rk4_step(lambda b: orc.geodesic_rhs(ost, b), y, hh)
        wall = time.perf_counter() - t0
        cpu = dict(value=20_000 * 50 / wall, unit="ray-steps/s", cores=1, kind="port",
                   sample="20,000 of the benchmark's rays x 50 RK4 steps, NumPy oracle rk4_step",
                   ns_per_ray_step=wall / (20_000 * 50) * 1e9)
        ref = sample.copy()
        for _ in range(50):
            ref = orc.rk4_step(lambda b: orc.geodesic_rhs(ost, b), ref, np.full(4096, 0.05))
        out["precise_equals_oracle_bitwise"] = bool(np.array_equal(got, ref))
    out["cpu_baseline"] = cpu
    out["reference_published_ns_per_ray_step"] = {
        "numpy_fp64_cpu": [1376, 1700], "cupy_fp64_rtx3070": 242.2}
    return out


def run_offline(args) -> dict:
    """The offline renderer's hero frame, tracing only: 1920 x 1200 at 2 x 2
    supersampling = 9.2 M rays, the renderer's defaults (a = 0.9, camera at
    26 M, inclination 82 deg, FOV 70 deg, rtol 1e-9, retrace at 1e-11 with
    4 x 12000 steps, disk from the ISCO to 18 M), both passes timed with
    HIP events (rays generated on the device beforehand).  The reference
    publishes 1645-1669 s for this frame on an RTX 3070 with CuPy, tracing
    plus CPU emission (docs/DESIGN.md:466-472, 503-504)."""
    import torch

    from blackhorizon_amd import _lib, kerr, offline

    _lib.require_device()
    w, h = (args.offline_size, args.offline_size * 5 // 8)
    settings = offline.OfflineSettings()
    cam = offline.FlyCamera.from_orbit(26.0, 82.0)
    st = kerr.KerrSpacetime(1.0, settings.spin)
    inner = offline.disk_inner_radius(st)
    outer = max(settings.disk_outer_radius, inner + 1.0)
    esc = 1.3 * max(cam.distance_from_origin, outer)
    s0 = offline.camera_rays(st, cam, w, h, settings)
    order = torch.from_numpy(offline.pixel_tile_order(w, h, settings.supersample, 4))
    cur = torch.cuda.current_stream()
    ms = []
    for k in range(args.warmup + args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(cur)
        res, retrace = offline.trace_tile_retrace(st, s0, esc, settings, (inner, outer),
                                                  mode=args.mode, ray_order=order)
        e1.record(cur)
        torch.cuda.synchronize()
        if k >= args.warmup:
            ms.append(e0.elapsed_time(e1))
        print(f"[bench] offline pass {k} done", file=sys.stderr, flush=True)
    t = float(np.median(ms))
    n = s0.shape[0]
    att1 = int(res.attempts.sum().item())
    steps1 = int(res.steps.sum().item())
    n_re = int(retrace.sum().item())
    counts = torch.bincount(res.status.long(), minlength=7).tolist()
    out = dict(rays=n, width=w, height=h, supersample=settings.supersample,
               mode=args.mode, ms_trace=t, first_pass_attempts=att1, first_pass_ray_steps=steps1,
               retraced_rays=n_re, status_counts=counts,
               first_pass_pct_fp64_peak_lower_bound=100.0 * att1 * FLOP_PER_ATTEMPT
               / (t * 1e-3) / (FP64_VECTOR_PEAK_TFLOPS * 1e12),
               reference_published_s=[1645, 1669])
    if not args.no_cpu:
        from oracle import bh_oracle as orc
        from oracle import offline_oracle as off

        rng = np.random.default_rng(5)
        idx = np.sort(rng.choice(n, size=1500, replace=False))
        ost = orc.Kerr(spin=settings.spin)
        dirs = off.subpixel_directions(off.FlyCamera.from_orbit(26.0, 82.0), w, h,
                                       settings.fov_degrees, settings.supersample)[idx]
        pos = np.tile(cam.position, (len(idx), 1))
        y0 = orc.build_state(pos, orc.null_momentum_from_velocity(ost, pos, dirs))
        t0 = time.perf_counter()
        r, _ = off.trace_tile_retrace(ost, y0, esc, (inner, outer), settings.rtol,
                                      settings.refine_rtol, settings.max_steps)
        wall = time.perf_counter() - t0
        out["cpu_baseline"] = dict(
            value=int(r.steps.sum()) / wall, unit="first-pass ray-steps/s", cores=1, kind="port",
            sample="1500 seeded rays of the frame, NumPy oracle of the offline loop, both passes",
            s_per_ray=wall / len(idx), projected_frame_s=wall / len(idx) * n)
    return out


def run_shadowfit(args) -> dict:
    """512 x 512 differentiable shadow fit at a = 0.7, inclination 45 deg,
    r_obs = 40 M, FOV 30 deg, escape sphere 45 M, rtol 1e-8, checkpoints
    every 32 accepted steps; target traced at a = 0.75, inclination 50 deg.
    Loss: squared angle of the escape directions on the escape sphere for
    rays escaped in both, of the plunge directions on the scoring sphere
    for rays captured in both, the squared capture margin for the others.
    Times the plain forward trace, the recording forward, the adjoint op
    (Recorded.backward for the loss's terminal adjoint) and the whole fit
    step (camera rays, recording trace, loss and autograd gradients)."""
    import math

    import torch

    from blackhorizon_amd import _lib, adjoint, tracer

    _lib.require_device()
    size = args.fit_size
    r_plus = lambda a: 1.0 + math.sqrt(1.0 - a * a)  # noqa: E731
    cfg = adjoint.TraceSettings(escape_radius=45.0, rtol=1e-8, atol=1e-10,
                                capture_margin=0.01 / r_plus(0.7), K=32, mode=args.mode)
    cfg_t = adjoint.TraceSettings(escape_radius=45.0, rtol=1e-8, atol=1e-10,
                                  capture_margin=0.01 / r_plus(0.75), K=32, mode=args.mode)
    with torch.no_grad():
        tgt = adjoint.camera_state(0.75, 50.0, 40.0, 30.0, size, size)
        Yt, st_t = adjoint.trace_terminal(tgt, 0.75, cfg_t)
        dirs_t = adjoint.terminal_directions(Yt, 0.75)
    p = [torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True)
         for v in (0.7, 45.0, 40.0)]
    y0 = adjoint.camera_state(p[0], p[1], p[2], 30.0, size, size)
    cur = torch.cuda.current_stream()
    # plain forward on the same rays (the reference-equivalent trace)
    st = tracer.KerrSpacetime(1.0, 0.7)

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # the terminal adjoint of the fit loss, for timing the adjoint op alone
    Y, status = adjoint.trace_terminal(y0, p[0], cfg)
    loss = adjoint.shadow_loss(Y, status, p[0], dirs_t, st_t, state0=y0)
    gY = torch.autograd.grad(loss, Y, retain_graph=True)[0].detach()
    y0d = y0.detach().contiguous()
    times = {"plain": [], "record": [], "adjoint": [], "fit_step": []}
    for k in range(args.warmup + args.steps):
        e = [ev() for _ in range(6)]
        torch.cuda.synchronize()
        e[0].record(cur)
        res = tracer.trace_rays(st, y0d, 45.0, rtol=1e-8, atol=1e-10,
                                capture_margin=cfg.capture_margin, mode=args.mode)
        e[1].record(cur)
        rec = adjoint.Recorded(y0d, 0.7, cfg)
        e[2].record(cur)
        rec.backward(gY)
        e[3].record(cur)
        # the whole differentiable fit step: camera rays, trace, loss, gradients
        y0 = adjoint.camera_state(p[0], p[1], p[2], 30.0, size, size)
        Y, status = adjoint.trace_terminal(y0, p[0], cfg)
        loss = adjoint.shadow_loss(Y, status, p[0], dirs_t, st_t, state0=y0)
        g = torch.autograd.grad(loss, p)
        e[4].record(cur)
        torch.cuda.synchronize()
        if k >= args.warmup:
            times["plain"].append(e[0].elapsed_time(e[1]))
            times["record"].append(e[1].elapsed_time(e[2]))
            times["adjoint"].append(e[2].elapsed_time(e[3]))
            times["fit_step"].append(e[3].elapsed_time(e[4]))
        del rec
        print(f"[bench] shadowfit pass {k} done", file=sys.stderr, flush=True)
    med = {k: float(np.median(v)) for k, v in times.items()}
    steps = int(res.steps.sum().item())
    profiled = {}
    if args.torch_profile:
        # one more fit step under the PyTorch profiler: where its device time goes
        from torch.profiler import ProfilerActivity, profile

        torch.cuda.synchronize()
        with profile(activities=[ProfilerActivity.CPU, ProfilerActivity.CUDA]) as prof:
            t0 = time.perf_counter()
            y0 = adjoint.camera_state(p[0], p[1], p[2], 30.0, size, size)
            Y, status = adjoint.trace_terminal(y0, p[0], cfg)
            loss = adjoint.shadow_loss(Y, status, p[0], dirs_t, st_t, state0=y0)
            g = torch.autograd.grad(loss, p)
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t0) * 1e3
        kern = [e for e in prof.events() if e.device_type.name == "CUDA"]
        busy = sum(e.device_time for e in kern) / 1e3
        tracer_ms = sum(e.device_time for e in kern if e.name.startswith(("bh::", "void bh::"))) / 1e3
        print(prof.key_averages().table(sort_by="cuda_time_total", row_limit=25,
                                        max_name_column_width=70), file=sys.stderr, flush=True)
        profiled = dict(profiled_fit_step=dict(wall_ms=wall, device_busy_ms=busy,
                                               tracer_kernel_ms=tracer_ms,
                                               device_launches=len(kern)))
    return dict(**profiled, rays=size * size, ray_steps=steps, loss=float(loss.detach()),
                grad_spin=float(g[0]), grad_inclination=float(g[1]), grad_distance=float(g[2]),
                ms_plain_forward=med["plain"], ms_recording_forward=med["record"],
                ms_adjoint=med["adjoint"], ms_fit_step=med["fit_step"],
                adjoint_over_forward=med["adjoint"] / med["plain"],
                forward_plus_adjoint_over_forward=(med["record"] + med["adjoint"]) / med["plain"],
                status_counts=torch.bincount(status.long(), minlength=7).tolist(),
                peak_device_gb=torch.cuda.max_memory_allocated() / 1e9)


def run_shader(args) -> dict:
    """The real-time shader's fixed-step algorithm (realtime/reference.py:
    41-188) over a 1920x1080 frame per quality preset, a = 0.9, camera at
    30 M and inclination 80 deg, FOV 60 deg, equatorial disk from the ISCO to
    18 M, the bit-exact build; plus the image stage at the same size: camera
    states built on the device and the Stage-1 shading of a traced frame.
    HIP-event times over the timed passes (rays resident on the device)."""
    import torch

    from blackhorizon_amd import _lib, imaging, kerr, offline, realtime, tracer
    from blackhorizon_amd.camera import PinholeCamera

    _lib.require_device()
    w, h = 1920, 1080
    st = kerr.KerrSpacetime(1.0, 0.9)
    cam = PinholeCamera.from_orbit(30.0, 80.0, fov_degrees=60.0, width=w, height=h)
    s0 = cam.device_states(st)
    disk = (offline.disk_inner_radius(st), 18.0)
    cur = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(cur)
        for _ in range(args.steps):
            out = fn()
        e1.record(cur)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps, out

    presets = {}
    for preset in realtime.QualityPreset:
        cfg = realtime.RenderSettings().apply_preset(preset)
        esc = cfg.effective_escape_radius(30.0)
        ms, res = timed(lambda: realtime.trace_like_shader(st, s0, cfg, esc, disk))
        counts = torch.bincount(res.status.long(), minlength=7).tolist()
        presets[preset.value] = dict(max_steps=cfg.max_steps, step_scale=cfg.step_scale,
                                     ms=ms, rays_per_s=w * h / (ms * 1e-3),
                                     status_counts=counts)
        print(f"[bench] shader preset {preset.value}: {ms:.2f} ms", file=sys.stderr, flush=True)
    ms_cam, _ = timed(lambda: cam.device_states(st))
    traced = tracer.trace_rays(st, s0, 60.0, rtol=1e-8, atol=1e-10)
    ms_shade, _ = timed(lambda: imaging.render_image(st, traced, w, h))
    return dict(rays=w * h, presets=presets, ms_camera_states=ms_cam,
                ms_render_image_incl_host_copy=ms_shade)


def run_orbits(args) -> dict:
    """65,536 near-circular timelike orbits around a = 0.9 in 16 uneven jobs,
    integrated to proper time 1e4 M at rtol 1e-10 in one persistent batched
    launch (--mode, "auto" = the precise build at this rtol); the same launch
    in the fast build, with its largest final-state difference from the
    first; E / L_z / H drift; adjoint of the final positions w.r.t. spin and
    initial momenta."""
    import torch

    from blackhorizon_amd import _lib, adjoint, batch, geodesics, kerr, tracer, workloads

    _lib.require_device()
    jobs = workloads.orbit_batches(n_total=args.orbits, lambda_end=args.proper_time)
    st = kerr.KerrSpacetime(1.0, 0.9)
    tjobs, states = [], []
    for j in jobs:
        pos = torch.from_numpy(j.positions).cuda()
        mom = geodesics.timelike_momentum(st, pos, torch.from_numpy(j.velocities).cuda())
        s0 = geodesics.build_state(pos, mom)
        states.append(s0)
        tjobs.append(batch.TraceJob(st, s0, j.escape_radius, j.max_steps, j.rtol, j.atol,
                                    j.initial_step, j.max_step, j.min_step, j.capture_margin,
                                    j.lambda_end))
    cur = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(p):
        for _ in range(args.warmup):
            p.launch()
        torch.cuda.synchronize()
        e0.record(cur)
        for _ in range(args.steps):
            p.launch()
        e1.record(cur)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    # one lane per orbit, then a DPP quad of lanes per orbit (the precise
    # build's quad tracer: the same results, each attempt's latency shorter)
    plan = batch.BatchPlan(tjobs, mode=args.mode)
    ms_one_lane = timed(plan)
    res = plan.results(as_numpy=False)
    del plan
    print("[bench] orbits forward (one lane per orbit) done", file=sys.stderr, flush=True)
    ms_quad, quad_identical = None, None
    if all(tracer.resolve_mode(args.mode, j.rtol) == "precise" for j in tjobs):
        qplan = batch.BatchPlan(tjobs, mode=args.mode, tune=tracer.Tune(lanes_per_ray=4))
        ms_quad = timed(qplan)
        qres = qplan.results(as_numpy=False)
        del qplan
        quad_identical = all(
            torch.equal(a.states.view(torch.int64), b.states.view(torch.int64))
            and torch.equal(a.status, b.status) and torch.equal(a.steps, b.steps)
            and torch.equal(a.attempts, b.attempts) and a.iterations == b.iterations
            for a, b in zip(res, qres))
        del qres
        print("[bench] orbits forward (quad per orbit) done", file=sys.stderr, flush=True)
    ms = ms_quad if ms_quad is not None else ms_one_lane
    # the same orbits in the fast build: every attempt's latency is about
    # half the precise build's; its final states are compared with the
    # precise ones below
    fplan = batch.BatchPlan(tjobs, mode="fast")
    for _ in range(args.warmup):
        fplan.launch()
    torch.cuda.synchronize()
    e0.record(cur)
    for _ in range(args.steps):
        fplan.launch()
    e1.record(cur)
    torch.cuda.synchronize()
    ms_fast = e0.elapsed_time(e1) / args.steps
    fres = fplan.results(as_numpy=False)
    del fplan
    print("[bench] orbits fast-build forward done", file=sys.stderr, flush=True)
    s0 = torch.cat(states)
    s1 = torch.cat([r.states for r in res])
    status = torch.cat([r.status for r in res])
    steps = int(sum(int(r.steps.sum().item()) for r in res))
    per_orbit = torch.cat([r.steps for r in res]).double()
    att = torch.cat([r.attempts for r in res]).double()
    d0 = geodesics._diagnostics(st, s0, ("ham", "energy", "lz"))
    d1 = geodesics._diagnostics(st, s1, ("ham", "energy", "lz"))
    ok = status == 6
    f1 = torch.cat([r.states for r in fres])
    fstatus = torch.cat([r.status for r in fres])
    both = ok & (fstatus == 6)
    # block-relative difference of the final states (t, position, p_t,
    # momentum; each relative to max(norm, 1)), fast vs precise
    fdiff = torch.zeros(int(both.sum().item()), dtype=torch.float64, device=s1.device)
    for lo, hi in ((0, 1), (1, 4), (4, 5), (5, 8)):
        a_, b_ = f1[both][:, lo:hi], s1[both][:, lo:hi]
        fdiff = torch.maximum(fdiff, (a_ - b_).norm(dim=1) / b_.norm(dim=1).clamp(min=1.0))
    dlz = ((d1["lz"] - d0["lz"]).abs() / d0["lz"].abs())[ok]
    dh = (d1["ham"] - d0["ham"]).abs()[ok]
    out = dict(orbits=int(s0.shape[0]), jobs=len(jobs), proper_time=args.proper_time,
               ray_steps=steps, ms_forward=ms, ray_steps_per_s=steps / (ms * 1e-3),
               ms_forward_one_lane=ms_one_lane, ms_forward_quad=ms_quad,
               quad_results_identical=quad_identical,
               ms_forward_fast_build=ms_fast,
               fast_build_statuses_identical=bool(torch.equal(fstatus, status)),
               fast_vs_precise_state_diff_max=float(fdiff.max()),
               fast_vs_precise_state_diff_median=float(fdiff.median()),
               status_counts=torch.bincount(status.long(), minlength=7).tolist(),
               energy_drift_max=float((d1["energy"] - d0["energy"]).abs().max()),
               lz_rel_drift_max=float(dlz.max()), lz_rel_drift_median=float(dlz.median()),
               ham_drift_max=float(dh.max()), ham_drift_median=float(dh.median()),
               # the launch cannot end before its longest orbit: attempts of
               # that orbit against the mean show how much of the time is tail
               steps_per_orbit={"median": float(per_orbit.median()),
                                "p99": float(torch.quantile(per_orbit, 0.99)),
                                "max": float(per_orbit.max())},
               attempts_total=int(att.sum().item()), attempts_max=int(att.max().item()),
               ns_per_attempt_of_longest=ms * 1e6 / float(att.max()))
    # adjoint of the final positions (one recorded launch over all orbits)
    cfg = adjoint.TraceSettings(escape_radius=1000.0, max_steps=10_000_000, rtol=1e-10,
                                atol=1e-12, initial_step=0.05, min_step=1e-14, K=32,
                                lambda_end=args.proper_time, mode=args.mode)
    sg = s0.clone().requires_grad_(True)
    a = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.randn((sg.shape[0], 3), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(3))
    # warm-up on a few orbits over a short proper time: loads the kernels
    # the timed pass uses (first launches of a module cost milliseconds)
    sw = s0[:256].clone().requires_grad_(True)
    Yw, _ = adjoint.trace_terminal(sw, a, dataclasses.replace(cfg, lambda_end=50.0))
    (Yw[:, 1:4] * w[:256]).sum().backward()
    a.grad = None
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record(cur)
    Y, st_ad = adjoint.trace_terminal(sg, a, cfg)
    e[1].record(cur)
    (Y[:, 1:4] * w).sum().backward()
    e[2].record(cur)
    torch.cuda.synchronize()
    out.update(ms_recording_forward=e[0].elapsed_time(e[1]), ms_adjoint=e[1].elapsed_time(e[2]),
               adjoint_over_forward=e[1].elapsed_time(e[2]) / ms,
               # the reverse pass runs one lane per orbit: against the
               # forward of the same lane layout
               adjoint_over_one_lane_forward=e[1].elapsed_time(e[2]) / ms_one_lane,
               grad_spin=float(a.grad),
               grad_momentum_norm_median=float(sg.grad[:, 5:8].norm(dim=1).median()),
               grad_finite=bool(torch.isfinite(sg.grad).all()))
    return out


def run_fit(args) -> dict:
    """Spin and inclination fit (adjoint.fit_shadow) of a --fit-image^2 camera
    (default 512^2, the shadow-fit configuration) at r_obs = 40 M, FOV 30
    deg, from (a, inc) = (0.7, 45 deg) to a target image traced at (0.75, 50
    deg): Adam on the shadow-fit loss with the adjoint's gradients, rays near
    the shadow's edge faded by their capture margin (edge scale 0.05), 60
    iterations, step sizes (0.005, 0.5 deg) decayed by 0.93 per iteration."""
    import math

    import torch

    from blackhorizon_amd import _lib, adjoint

    _lib.require_device()
    size = args.fit_image
    rp = 1.0 + math.sqrt(1.0 - 0.75 ** 2)
    with torch.no_grad():
        tgt = adjoint.camera_state(0.75, 50.0, 40.0, 30.0, size, size)
        Yt, st_t = adjoint.trace_terminal(tgt, 0.75, adjoint.TraceSettings(
            escape_radius=45.0, rtol=1e-8, atol=1e-10, capture_margin=0.01 / rp, K=32))
        dirs_t = adjoint.terminal_directions(Yt, 0.75)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = adjoint.fit_shadow(dirs_t, st_t, 0.7, 45.0, 40.0, 30.0, size, size,
                              edge_scale=0.05)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    first, last = hist[0], hist[-1]
    return dict(rays=size * size, iterations=len(hist) - 1, seconds=wall,
                ms_per_iteration=wall * 1e3 / (len(hist) - 1),
                start=dict(spin=first.spin, inclination=first.inclination, loss=first.loss),
                final=dict(spin=last.spin, inclination=last.inclination, loss=last.loss,
                           mismatched=last.mismatched),
                target=dict(spin=0.75, inclination=50.0), edge_scale=0.05,
                trajectory=[[h.iteration, round(h.spin, 6), round(h.inclination, 4),
                             round(h.loss, 4)] for h in hist[::5]])


def run_dense(args) -> dict:
    """integrate_dense (integrators.py:127-174) on the device: --orbits
    near-circular orbits around a = 0.9 outside r = 15 M (2% velocity
    spread, so no orbit plunges and sinks the shared step), one batch-shared
    adaptive step over proper time --dense-lambda at rtol 1e-10, precise
    build, the whole trajectory kept on the device.  Every attempt runs all
    rays, so FLOP = attempts x rays x 1321."""
    import torch

    from blackhorizon_amd import _lib, geodesics, integrators, kerr, workloads

    _lib.require_device()
    st = kerr.KerrSpacetime(1.0, 0.9)
    p, v = workloads.near_circular_orbits(args.orbits, seed=3, r_min=15.0, spread=0.02)
    pos = torch.from_numpy(p).cuda()
    s0 = geodesics.build_state(pos, geodesics.timelike_momentum(st, pos,
                                                                torch.from_numpy(v).cuda()))
    g = integrators.GeodesicRHS(st, mode="precise")
    times = []
    traj = None
    for k in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        traj = integrators.integrate_dense(g, s0, args.dense_lambda, rtol=1e-10, chunk=1024)
        torch.cuda.synchronize()
        if k >= args.warmup:
            times.append(time.perf_counter() - t0)
        print(f"[bench] dense pass {k} done", file=sys.stderr, flush=True)
    n = s0.shape[0]
    accepted = len(traj) - 1
    ms = float(np.median(times)) * 1e3
    return dict(orbits=n, proper_time=args.dense_lambda, accepted_steps=accepted,
                ray_steps=accepted * n, ms_median=ms, ray_steps_per_s=accepted * n / (ms * 1e-3),
                us_per_accepted_step=ms * 1e3 / max(accepted, 1))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hd",
                    choices=["hd", "hd09", "small", "uneven", "shadowfit", "orbits", "rk4",
                             "offline", "shader", "dense", "fit"])
    ap.add_argument("--dense-lambda", type=float, default=100.0,
                    help="dense: proper time of the batch-shared integration")
    ap.add_argument("--orbits", type=int, default=65536, help="orbits: number of orbits")
    ap.add_argument("--proper-time", type=float, default=1.0e4, help="orbits: proper time")
    ap.add_argument("--fit-size", type=int, default=512, help="shadowfit: image side")
    ap.add_argument("--torch-profile", action="store_true",
                    help="shadowfit: one more fit step under the PyTorch profiler")
    ap.add_argument("--fit-image", type=int, default=512, help="fit: image side")
    ap.add_argument("--offline-size", type=int, default=1920,
                    help="offline: frame width (height = 5/8 of it)")
    ap.add_argument("--batch-order", default="longest_first", choices=["longest_first", "given"],
                    help="uneven: queue order of the batched launch (per-job: row order)")
    ap.add_argument("--serial-groups", action="store_true",
                    help="uneven: run the fast and precise batches one after the other")
    ap.add_argument("--mode", default="auto", choices=["auto", "fast", "precise"])
    ap.add_argument("--tune", default="auto",
                    help='launch shape: "auto" (autotuner, untimed, before the sweep), '
                         '"default" (256,1,1) or "block,rays_per_lane,refill_min"')
    ap.add_argument("--order", default="center8",
                    help="ray hand-out order: row (pixel order), tile<k> (k x k pixel tiles, "
                         "row-major) or center<k> (k x k tiles, nearest the image centre first)")
    ap.add_argument("--cpu-rays-per-frame", type=int, default=12288)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="hd/hd09: skip the adjoint/forward measurement (512^2 shadow fit)")
    ap.add_argument("--no-targets", action="store_true",
                    help="hd: skip the uneven-mix, timelike-orbit and frame-copy measurements")
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build targets one GPU: the tracer shards rays with no "
                                   "exchange, but only the single-GPU path is implemented and "
                                   "measured; run with --gpus 1"}))
        return 2

    if args.workload == "fit":
        res = run_fit(args)
        print(json.dumps({"metric": "spin and inclination fit to a synthetic shadow",
                          "workload": "fit", **res}), flush=True)
        return 0
    if args.workload == "dense":
        res = run_dense(args)
        print(json.dumps({"metric": "integrate_dense (batch-shared step) ray-steps/s",
                          "workload": "dense", **res}), flush=True)
        return 0
    if args.workload == "orbits":
        res = run_orbits(args)
        print(json.dumps({"metric": METRIC, "workload": "timelike orbits", "mode": args.mode,
                          **res}), flush=True)
        return 0
    if args.workload == "shader":
        res = run_shader(args)
        print(json.dumps({"metric": "fixed-step shader algorithm and image stage, 1920x1080",
                          "workload": "shader", **res}), flush=True)
        return 0
    if args.workload == "offline":
        res = run_offline(args)
        print(json.dumps({"metric": "offline hero frame tracing time", "workload": "offline",
                          **res}), flush=True)
        return 0
    if args.workload == "rk4":
        res = run_rk4(args)
        print(json.dumps({"metric": "ns per ray-step, fixed-step RK4 (reference benchmark)",
                          "workload": "rk4", **res}), flush=True)
        return 0
    if args.workload == "shadowfit":
        res = run_shadowfit(args)
        print(json.dumps({"metric": METRIC, "workload": "shadow fit", "mode": args.mode,
                          **res}), flush=True)
        return 0
    if args.workload == "uneven":
        res = run_uneven(args)
        print(json.dumps({"metric": METRIC, "workload": "uneven 64-job mix", "mode": args.mode,
                          **res}), flush=True)
        return 0

    jobs = frames(args.workload)
    cpu = cpu_par = None
    if not args.no_cpu:
        cpu = cpu_baseline(jobs, args.cpu_rays_per_frame, processes=1)
        p = min(len(os.sched_getaffinity(0)), 16)
        if p > 1:
            cpu_par = cpu_baseline(jobs, args.cpu_rays_per_frame, processes=p)
        print(f"[bench] cpu baseline done: {cpu['ns_per_ray_step']:.0f} ns/ray-step (1 core)",
              file=sys.stderr, flush=True)

    import torch

    from blackhorizon_amd import _lib, tracer

    _lib.require_device()
    prepared = prepare_gpu(jobs, args.mode, args.order)
    # tune on the a = 0.9, 17 deg frame when present (else the first)
    tune_plans = [jp for jp in prepared if jp[0].spin == 0.9] or prepared
    tune, tune_info = pick_tune(args, tune_plans, f"{args.workload}-a0.9-{args.order}",
                                workload=prepared)
    cur = torch.cuda.current_stream()

    for w in range(args.warmup):
        for job, plan in prepared:
            plan.launch(tune)
        torch.cuda.synchronize()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        for job, plan in prepared:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(cur)
            plan.launch(tune)
            e1.record(cur)
            events.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    kern_ms = [a.elapsed_time(b) for a, b in events]
    per_frame = []
    steps_sweep = attempts_sweep = 0
    for i, (job, plan) in enumerate(prepared):
        s = int(plan.steps.sum().item())
        a = int(plan.attempts.sum().item())
        st = torch.bincount(plan.status.long(), minlength=7).tolist()
        ms = [kern_ms[k * len(prepared) + i] for k in range(args.steps)]
        per_frame.append(dict(spin=job.spin, inclination=job.inclination_degrees,
                              lane_efficiency=round(lane_efficiency(plan, tune.rays_per_lane), 4),
                              rays=job.n_rays, ray_steps=s, attempts=a,
                              iterations=int(plan.iters.item()),
                              kernel_ms=round(float(np.mean(ms)), 3),
                              kernel_ms_median=round(float(np.median(ms)), 4),
                              ns_per_ray_step=float(np.mean(ms)) * 1e6 / max(s, 1),
                              pct_fp64_peak=100.0 * a * FLOP_PER_ATTEMPT
                              / (float(np.mean(ms)) * 1e-3) / (FP64_VECTOR_PEAK_TFLOPS * 1e12),
                              status_counts=st[:7]))
        steps_sweep += s
        attempts_sweep += a
    targets = None
    if args.workload == "hd" and not args.no_targets:
        # the north star's other targets, outside the timed sweep: the
        # uneven 64-job mix (one persistent batched launch vs one plain
        # launch per job, both autotuned), the 65,536-orbit timelike
        # workload, and one frame's ray-generation and copy costs
        io_job = next(j for j in jobs if j.spin == 0.9)
        del prepared
        torch.cuda.empty_cache()
        targets = {"frame_io": frame_io(io_job, args.mode, tune)}
        print("[bench] frame ray generation and copies measured", file=sys.stderr, flush=True)
        sub = argparse.Namespace(**{**vars(args), "steps": 3, "warmup": 1, "tune_reps": 1})
        un = run_uneven(sub)
        targets["uneven"] = {k: un[k] for k in (
            "speedup", "ms_batched", "ms_per_job", "speedup_graph", "ms_batched_graph",
            "ms_per_job_graph", "batched_ray_steps_per_s", "per_job_ray_steps_per_s",
            "pct_fp64_vector_peak_batched", "rays", "ray_steps", "per_ray_steps_identical",
            "jobs_per_mode", "tune", "tune_per_job")}
        print("[bench] uneven mix measured", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        orb = run_orbits(argparse.Namespace(**{**vars(args), "steps": 3, "warmup": 1,
                                               "orbits": 65536, "proper_time": 1.0e4}))
        targets["orbits"] = {k: orb[k] for k in (
            "orbits", "jobs", "proper_time", "ray_steps", "ms_forward", "ray_steps_per_s",
            "ms_forward_one_lane", "ms_forward_quad", "quad_results_identical",
            "ms_forward_fast_build", "fast_build_statuses_identical",
            "fast_vs_precise_state_diff_max",
            "status_counts", "energy_drift_max", "lz_rel_drift_max", "lz_rel_drift_median",
            "ham_drift_max", "ham_drift_median", "steps_per_orbit", "ms_recording_forward",
            "ms_adjoint", "adjoint_over_forward", "adjoint_over_one_lane_forward",
            "grad_finite")}
        print("[bench] timelike orbits measured", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        fit = run_fit(argparse.Namespace(**{**vars(args), "fit_image": 512}))
        targets["fit"] = {k: fit[k] for k in (
            "rays", "iterations", "ms_per_iteration", "start", "final", "target", "edge_scale")}
        print("[bench] 512^2 spin/inclination fit done", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    adj = None
    if not args.no_adjoint and args.workload in ("hd", "hd09"):
        # the metric's third part: adjoint op time over plain forward time on
        # the 512^2 shadow fit (a few seconds, outside the timed sweep)
        fit_args = argparse.Namespace(**{**vars(args), "steps": 3, "warmup": 1})
        adj = run_shadowfit(fit_args)
        print("[bench] adjoint/forward measured", file=sys.stderr, flush=True)
    total_kernel_s = sum(kern_ms) * 1e-3
    sweep_median_ms = float(sum(f["kernel_ms_median"] for f in per_frame))
    value = steps_sweep * args.steps / wall
    flop = attempts_sweep * args.steps * FLOP_PER_ATTEMPT
    achieved_tflops = flop / total_kernel_s / 1e12
    props = torch.cuda.get_device_properties(0)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": {"hd": "1920x1080 sweep, 8 frames: a in {0,0.5,0.9,0.998} x inc "
                               "{17,80} deg, r_obs 50 M, FOV 25 deg, rtol 1e-8, atol 1e-10",
                         "hd09": "1920x1080, a = 0.9, inc {17,80} deg, r_obs 50 M, FOV 25 deg, "
                                 "rtol 1e-8, atol 1e-10",
                         "small": "160x120, a = 0.9, r_obs 30 M, inc 60 deg, FOV 30 deg, "
                                  "rtol 1e-8, atol 1e-10"}[args.workload],
            "rays_per_step": int(sum(j.n_rays for j in jobs)),
            "launches_per_step": len(jobs),
            "mode": args.mode,
            "tune": {"block": tune.block_size, "rays_per_lane": tune.rays_per_lane,
                     "refill_min": tune.refill_min, "autotune": tune_info},
            "order": args.order,
        },
        "ns_per_ray_step": wall / (steps_sweep * args.steps) * 1e9,
        "attempts_per_s": attempts_sweep * args.steps / wall,
        "pct_fp64_vector_peak": 100.0 * achieved_tflops / FP64_VECTOR_PEAK_TFLOPS,
        "adjoint_over_forward": None if adj is None else adj["adjoint_over_forward"],
        "batched_over_per_job": None if targets is None else targets["uneven"]["speedup"],
        "orbits_forward_ms": None if targets is None else targets["orbits"]["ms_forward"],
        # per frame the median of its timed launches, summed over the sweep
        "sweep_kernel_ms_median": sweep_median_ms,
        "pct_fp64_vector_peak_median": 100.0 * attempts_sweep * FLOP_PER_ATTEMPT / (
            sweep_median_ms * 1e-3) / (FP64_VECTOR_PEAK_TFLOPS * 1e12),
        "targets": targets,
        "adjoint": None if adj is None else {
            k: adj[k] for k in ("rays", "ray_steps", "ms_plain_forward", "ms_recording_forward",
                                "ms_adjoint", "ms_fit_step", "adjoint_over_forward")},
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": HD_TRAFFIC_BYTES_PER_LAUNCH if (
                args.workload == "hd"
                and tracer.resolve_mode(args.mode, jobs[0].rtol) == "fast") else None,
            "kernel": "trace_plain_kernel ("
                      + tracer.resolve_mode(args.mode, jobs[0].rtol) + " build)",
            "flop_per_launch": flop / len(events),
            "mean_launch_ms": float(np.mean(kern_ms)),
            "kernel_resources": _lib.kernel_attributes(
                "plain", tracer.resolve_mode(args.mode, jobs[0].rtol), tune.block_size),
        },
        "cpu_baseline": None if cpu is None else {
            k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "cpu_baseline_detail": cpu,
        "cpu_baseline_parallel": cpu_par,
        "gpu": {"name": props.name, "cus": props.multi_processor_count},
        "per_frame": per_frame,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
