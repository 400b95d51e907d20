"""GPU: the quad tracer (the four lanes of a DPP quad trace one ray, precise
build, tune lanes_per_ray = 4) gives every ray bit for bit what the
one-lane batched tracer gives it -- states compared as raw 64-bit patterns,
so signed zeros count -- on camera jobs of the uneven mix, on the failure
and cap edge cases, and on the whole 65,536-orbit timelike workload."""

import math
import time

import numpy as np
import pytest

from conftest import golden
from oracle import bh_oracle as orc
from blackhorizon_amd import workloads

pytestmark = pytest.mark.gpu

EDGE_CASES = ("default", "nonfinite", "max_steps_0", "max_steps_1", "max_steps_7",
              "min_step_high", "tight_small_h", "loose")


def _np(a):
    import torch

    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _assert_bits(a, b, label):
    sa, sb = _np(a.states), _np(b.states)
    assert sa.shape == sb.shape, label
    np.testing.assert_array_equal(sa.view(np.int64), sb.view(np.int64), err_msg=label)
    np.testing.assert_array_equal(_np(a.status), _np(b.status), err_msg=label)
    np.testing.assert_array_equal(_np(a.steps), _np(b.steps), err_msg=label)
    np.testing.assert_array_equal(_np(a.attempts), _np(b.attempts), err_msg=label)
    assert a.iterations == b.iterations, label


def _camera_jobs(hip, n_jobs, stride):
    from blackhorizon_amd import batch

    jobs = []
    for j in workloads.uneven_mix()[:n_jobs]:
        st = orc.Kerr(spin=j.spin)
        cam = orc.PinholeCamera.from_orbit(j.distance, j.inclination_degrees,
                                           fov_degrees=j.fov_degrees, width=j.width,
                                           height=j.height)
        s0 = orc.camera_state(st, cam)[::stride]
        jobs.append(batch.TraceJob(hip.KerrSpacetime(spin=j.spin), s0, j.escape_radius,
                                   max_steps=j.max_steps, rtol=j.rtol, atol=j.atol,
                                   capture_margin=j.capture_margin))
    return jobs


@pytest.mark.parametrize("shape", [(256, 1), (64, 1), (128, 4), (256, 16)])
def test_quad_equals_one_lane_on_uneven_jobs(hip, shape):
    """16 camera jobs of the uneven mix (every 9th ray: spins of both
    signs, distances 20-1000 M, rtol 1e-10..1e-6) in the precise build,
    several block sizes and refill thresholds."""
    from blackhorizon_amd import batch

    jobs = _camera_jobs(hip, 16, 9)
    one = batch.trace_batch(jobs, mode="precise")
    quad = batch.trace_batch(jobs, mode="precise",
                             tune=hip.Tune(block_size=shape[0], refill_min=shape[1],
                                           lanes_per_ray=4))
    for k, (a, b) in enumerate(zip(one, quad)):
        _assert_bits(b, a, f"job {k}, block {shape[0]}, refill {shape[1]}")


def test_quad_edge_cases(hip):
    """Non-finite rows (FAILED), max_steps 0 / 1 / 7 and the attempt cap
    (MAX_STEPS), h below min_step (FAILED), tiny and huge steps, a ray
    starting inside the horizon, and empty jobs, as one quad launch: every
    job equals the one-lane launch (which the parity tests hold to the
    reference's outcomes) and has the reference's iteration count."""
    from blackhorizon_amd import batch

    g = golden("edge_cases.npz")
    st = hip.KerrSpacetime(spin=0.9)
    jobs, refs = [batch.TraceJob(st, np.zeros((0, 8)), 40.0)], [None]
    for case in EDGE_CASES:
        pre = case + "_kw_"
        kw = {k[len(pre):]: float(g[k]) for k in g.files if k.startswith(pre)}
        if "max_steps" in kw:
            kw["max_steps"] = int(kw["max_steps"])
        jobs.append(batch.TraceJob(st, g[case + "_state0"], **kw))
        refs.append(case)
    one = batch.trace_batch(jobs, mode="precise")
    quad = batch.trace_batch(jobs, mode="precise", tune=hip.Tune(lanes_per_ray=4))
    for k, (a, b, case) in enumerate(zip(one, quad, refs)):
        _assert_bits(b, a, f"edge job {k} ({case})")
        if case is None:
            assert b.states.shape == (0, 8) and b.iterations == 0
            continue
        assert b.iterations == int(g[case + "_iterations"]), case


def test_quad_mixed_builds_fall_back_to_one_lane(hip):
    """lanes_per_ray 4 applies to the precise group of a plan; its fast
    group runs one lane per ray.  Results equal the one-lane plan's."""
    from blackhorizon_amd import batch

    jobs = _camera_jobs(hip, 8, 17)
    a = batch.BatchPlan(jobs, mode="auto")
    b = batch.BatchPlan(jobs, mode="auto", tune=hip.Tune(lanes_per_ray=4))
    assert {g.mode for g in b.groups} == {"fast", "precise"}
    a.launch()
    b.launch()
    for k, (x, y) in enumerate(zip(a.results(), b.results())):
        _assert_bits(y, x, f"job {k}")


def test_quad_orbit_workload_full_size(hip):
    """The 65,536-orbit timelike workload (16 jobs, a = 0.9, tau = 1e4 M,
    rtol 1e-10, precise): the quad launch equals the one-lane launch bit for
    bit on every orbit, and the launch takes less time."""
    import torch

    from blackhorizon_amd import batch, geodesics, kerr

    jobs = workloads.orbit_batches()
    st = kerr.KerrSpacetime(1.0, 0.9)
    tjobs = []
    for j in jobs:
        pos = torch.from_numpy(j.positions).cuda()
        mom = geodesics.timelike_momentum(st, pos, torch.from_numpy(j.velocities).cuda())
        tjobs.append(batch.TraceJob(st, geodesics.build_state(pos, mom), j.escape_radius,
                                    j.max_steps, j.rtol, j.atol, j.initial_step, j.max_step,
                                    j.min_step, j.capture_margin, j.lambda_end))
    times = {}
    res = {}
    for lanes in (1, 4):
        plan = batch.BatchPlan(tjobs, mode="precise", tune=hip.Tune(lanes_per_ray=lanes))
        plan.launch()  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.launch()
        torch.cuda.synchronize()
        times[lanes] = time.perf_counter() - t0
        res[lanes] = plan.results(as_numpy=False)
        del plan
    for k, (a, b) in enumerate(zip(res[1], res[4])):
        _assert_bits(b, a, f"orbit job {k}")
    print(f"orbit workload: one lane per orbit {times[1] * 1e3:.1f} ms, "
          f"a quad per orbit {times[4] * 1e3:.1f} ms")
    assert times[4] < times[1]


def test_integrate_rays_takes_the_quad_path_with_the_same_bits(hip):
    """integrate_rays of at most one wavefront per SIMD in the precise build
    runs on the quad tracer: 3,000 orbits to tau = 300 give what the plain
    one-lane launch (trace_rays with the same lambda_end, and
    integrate_rays with an explicit launch shape) gives, bit for bit."""
    import torch

    from blackhorizon_amd import geodesics, kerr, tracer

    pos, vel = workloads.near_circular_orbits(3000, seed=21)
    st = kerr.KerrSpacetime(1.0, 0.9)
    p = torch.from_numpy(pos).cuda()
    s0 = geodesics.build_state(p, geodesics.timelike_momentum(st, p, torch.from_numpy(vel).cuda()))
    kw = dict(rtol=1e-10, atol=1e-12, initial_step=0.05, min_step=1e-14, max_steps=10_000_000)
    quad = tracer.integrate_rays(st, s0, 300.0, **kw)
    plain = tracer.trace_rays(st, s0, math.inf, mode="precise", lambda_end=300.0, **kw)
    tuned = tracer.integrate_rays(st, s0, 300.0, tune=tracer.Tune(), **kw)
    for other in (plain, tuned):
        _assert_bits(quad, other, "integrate_rays")
    assert (_np(quad.status) == 6).mean() > 0.95
