"""GPU: launch shapes and HIP-graph replay change nothing in the results.

Every ray runs the same per-ray controller whatever the block size, rays
per lane or refill threshold, and a captured launch sequence replays the
same kernels on the same buffers, so results are compared bitwise."""

import numpy as np
import pytest
import torch

from oracle import bh_oracle as orc
from blackhorizon_amd import workloads

pytestmark = pytest.mark.gpu


def _small_rays(stride=5):
    job = workloads.small_frame()
    st = orc.Kerr(spin=job.spin)
    cam = orc.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees, width=job.width,
                                       height=job.height)
    return job, orc.camera_state(st, cam)[::stride]


def _plan(hip, job, s0, **kw):
    from blackhorizon_amd import tracer

    return tracer.TracePlan(hip.KerrSpacetime(spin=job.spin), torch.from_numpy(s0).cuda(),
                            job.escape_radius, job.max_steps, job.rtol, job.atol,
                            capture_margin=job.capture_margin, **kw)


def _snapshot(plan):
    return [t.clone() for t in (plan.soa_out, plan.status, plan.steps, plan.attempts,
                                plan.iters)]


def _assert_equal(a, b, label):
    for x, y in zip(a, b):
        assert torch.equal(x, y), label


def test_every_plain_candidate_gives_identical_results(hip):
    from blackhorizon_amd import autotune

    job, s0 = _small_rays()
    plan = _plan(hip, job, s0)
    plan.launch()
    ref = _snapshot(plan)
    for t in autotune.plain_candidates():
        plan.launch(t)
        _assert_equal(_snapshot(plan), ref, f"tune {t}")


def test_autotune_picks_a_candidate_and_caches_it(hip, tmp_path):
    from blackhorizon_amd import autotune

    job, s0 = _small_rays(stride=3)
    plan = _plan(hip, job, s0)
    cache = str(tmp_path / "tune.json")
    cands = autotune.plain_candidates()[:5]
    r1 = autotune.autotune(plan, "small", cache=cache, candidates=cands, reps=1)
    assert not r1.cached and r1.best in cands and len(r1.table) == 5
    assert plan.tune == r1.best
    r2 = autotune.autotune(plan, "small", cache=cache, candidates=cands)
    assert r2.cached and r2.best == r1.best
    # the tuned plan's results equal trace_rays' with the default shape
    plan.launch()
    res = plan.result()
    alone = hip.trace_rays(hip.KerrSpacetime(spin=job.spin), s0, job.escape_radius,
                           job.max_steps, job.rtol, job.atol, capture_margin=job.capture_margin)
    np.testing.assert_array_equal(res.states.cpu().numpy(), alone.states)
    np.testing.assert_array_equal(res.status.cpu().numpy(), alone.status)
    np.testing.assert_array_equal(res.steps.cpu().numpy(), alone.steps)


def test_plain_graph_replays_match_eager(hip):
    """Three replays, each after unrelated eager work on the same stream,
    each checked against the eager launch (also with a rewritten input)."""
    from blackhorizon_amd import graph

    job, s0 = _small_rays()
    plan = _plan(hip, job, s0, ray_order=np.random.default_rng(1).permutation(len(s0)))
    plan.launch()
    eager = _snapshot(plan)
    g = graph.TraceGraph(plan)
    for k in range(3):
        hip.trace_rays(hip.KerrSpacetime(spin=0.5), s0[:100], job.escape_radius)  # eager work
        plan.soa_out.fill_(float("nan"))
        plan.status.fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        _assert_equal(_snapshot(plan), eager, f"replay {k}")
    # inputs rewritten in place: the replay traces the new rays
    _, s1 = _small_rays(stride=7)
    m = min(len(s1), plan.n)
    other = _plan(hip, job, np.concatenate([s1[:m], s0[m:plan.n]]))
    other.launch()
    plan.soa_in.copy_(other.soa_in)
    g.replay()
    torch.cuda.synchronize()
    for x, y in zip(_snapshot(plan), _snapshot(other)):
        assert torch.equal(x, y), "rewritten input"


def test_batch_graph_replays_match_eager(hip):
    """Two arithmetic groups on two streams, captured as one graph."""
    from blackhorizon_amd import batch, graph

    mix = workloads.uneven_mix()[:6]
    jobs = []
    for j in mix:
        st = orc.Kerr(spin=j.spin)
        cam = orc.PinholeCamera.from_orbit(j.distance, j.inclination_degrees,
                                           fov_degrees=j.fov_degrees, width=j.width,
                                           height=j.height)
        jobs.append(batch.TraceJob(hip.KerrSpacetime(spin=j.spin),
                                   orc.camera_state(st, cam)[::29], j.escape_radius,
                                   max_steps=j.max_steps, rtol=j.rtol, atol=j.atol,
                                   capture_margin=j.capture_margin))
    plan = batch.BatchPlan(jobs, mode="auto")
    assert len(plan.groups) == 2
    plan.launch()
    eager = [_group_snapshot(gr) for gr in plan.groups]
    g = graph.TraceGraph(plan)
    for k in range(3):
        for gr in plan.groups:
            gr.soa_out.fill_(float("nan"))
            gr.steps.fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        for gr, e in zip(plan.groups, eager):
            _assert_equal(_group_snapshot(gr), e, f"replay {k}")


def _group_snapshot(gr):
    return [t.clone() for t in (gr.soa_out, gr.status, gr.steps, gr.attempts, gr.iters)]


def test_kernel_attributes_report_resources(hip):
    from blackhorizon_amd import _lib

    for kernel in ("plain", "batched", "record", "tile", "shader"):
        for mode in ("fast", "precise"):
            a = _lib.kernel_attributes(kernel, mode, 256)
            assert 32 <= a["vgprs"] <= 512, (kernel, mode, a)
            assert a["blocks_per_cu"] >= 1 and a["waves_per_simd"] >= 1
    a = _lib.kernel_attributes("adjoint", "fast", 64)
    assert a["lds_bytes"] == 20480 and a["waves_per_simd"] >= 1
    with pytest.raises(ValueError):
        _lib.kernel_attributes("adjoint", "fast", 256)
