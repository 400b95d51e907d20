"""GPU: the full-size workloads.  Whole 1920x1080 frames of the
large-trace configuration (2,073,600 rays per launch), all eight in the
auto (here fast) and in the precise build, checked through what does not
depend on the size:

* the device-built camera states at the golden file's 96 seeded pixels equal
  the oracle's bit for bit, and those pixels' results inside the full
  launch meet the parity bar against the reference's own traces of them
  (identical status and accepted steps, final states within 1e-7 or 10x
  the reference's host-to-host spread);
* the per-ray results of the whole frame are bitwise identical under
  another hand-out order and launch shape (row-major pixels, 64-thread
  blocks, in-wave refill) and a repeated launch;
* p_t is carried exactly (E conserved bit for bit), every ray ends
  captured or escaped, and the launch's iteration count is the largest
  per-ray attempt count.

The whole 64-job uneven mix is checked the same way (test_full_uneven_mix).
"""

import numpy as np
import pytest
import torch

from conftest import golden
from parity import assert_trace_parity, reference_spread
from blackhorizon_amd import workloads

pytestmark = pytest.mark.gpu

# every frame of the sweep in both builds
CASES = [(f, m) for m in ("auto", "precise") for f in range(8)]
_SPREAD = {}


def _spread(frame, job, g, kw):
    """The reference's own host-to-host spread on the sampled pixels (cached
    across modes)."""
    if frame not in _SPREAD:
        _SPREAD[frame] = reference_spread(job.spin, g[f"f{frame}_state0"],
                                          g[f"f{frame}_states"],
                                          escape_radius=job.escape_radius, **kw)
    return _SPREAD[frame]


@pytest.mark.parametrize("frame,mode", CASES)
def test_full_hd_frame(hip, frame, mode):
    from blackhorizon_amd.tracer import Tune

    g = golden("hd_subsample.npz")
    idx = g[f"f{frame}_index"]
    job = workloads.hd_frames()[frame]
    st = hip.KerrSpacetime(spin=job.spin)
    cam = hip.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees, width=job.width,
                                       height=job.height)
    s0 = cam.device_states(st)
    assert s0.shape == (job.width * job.height, 8)
    np.testing.assert_array_equal(s0[torch.from_numpy(idx).cuda()].cpu().numpy(),
                                  g[f"f{frame}_state0"])
    kw = dict(max_steps=job.max_steps, rtol=job.rtol, atol=job.atol,
              capture_margin=job.capture_margin)
    order = torch.from_numpy(cam.center_tile_order(8)).cuda()
    res = hip.trace_rays(st, s0, job.escape_radius, ray_order=order, mode=mode, **kw)

    # the sampled pixels of the full launch against the reference's traces
    sel = torch.from_numpy(idx).cuda()
    spread = _spread(frame, job, g, kw)
    assert_trace_parity(res.states[sel].cpu().numpy(), res.status[sel].cpu().numpy(),
                        res.steps[sel].cpu().numpy(), g[f"f{frame}_states"],
                        g[f"f{frame}_status"], g[f"f{frame}_steps"],
                        label=f"full frame {frame} sample", spread=spread)

    # the whole frame: bitwise independent of hand-out order and launch shape
    other = hip.trace_rays(st, s0, job.escape_radius, tune=Tune(64, 2, 16), mode=mode, **kw)
    again = hip.trace_rays(st, s0, job.escape_radius, ray_order=order, mode=mode, **kw)
    for r2 in (other, again):
        assert torch.equal(r2.status, res.status)
        assert torch.equal(r2.steps, res.steps)
        assert torch.equal(r2.attempts, res.attempts)
        assert torch.equal(r2.states.view(torch.int64), res.states.view(torch.int64))
        assert r2.iterations == res.iterations

    # conservation and termination
    assert torch.equal(res.states[:, 4].view(torch.int64), s0[:, 4].view(torch.int64))
    counts = torch.bincount(res.status.long(), minlength=7).cpu().numpy()
    assert counts[1] + counts[2] == job.width * job.height, counts
    assert counts[1] > 0 and counts[2] > 0
    assert int(res.attempts.max().item()) == res.iterations
    assert bool((res.attempts >= res.steps).all())


@pytest.mark.parametrize("mode", ["auto", "precise"])
def test_full_uneven_mix(hip, mode):
    """The whole 64-job uneven mix (2.19 M rays, spins of both signs,
    r_obs 20-1000 M, rtol 1e-10..1e-6; auto splits it between the two builds,
    precise runs every job bit-exact) in one batched plan
    with its default longest-first queue: every job's results equal a plain
    trace_rays launch of that job bit for bit, 12 rays of every job traced
    by the reference (uneven_all.npz) meet the parity bar inside the full
    launch, and p_t is carried exactly."""
    from blackhorizon_amd import batch

    g = golden("uneven_all.npz")
    mix = workloads.uneven_mix()
    jobs, spaces = [], []
    for j in mix:
        st = hip.KerrSpacetime(spin=j.spin)
        cam = hip.PinholeCamera.from_orbit(j.distance, j.inclination_degrees,
                                           fov_degrees=j.fov_degrees, width=j.width,
                                           height=j.height)
        spaces.append(st)
        jobs.append(batch.TraceJob(st, cam.device_states(st), j.escape_radius, j.max_steps,
                                   j.rtol, j.atol, j.initial_step, j.max_step, j.min_step,
                                   j.capture_margin))
    plan = batch.BatchPlan(jobs, mode=mode)
    assert len(plan.groups) == (2 if mode == "auto" else 1)
    plan.launch()
    res = plan.results()
    for k, (j, tj, r) in enumerate(zip(mix, jobs, res)):
        alone = hip.trace_rays(spaces[k], tj.state0, j.escape_radius, j.max_steps, j.rtol,
                               j.atol, j.initial_step, j.max_step, j.min_step, j.capture_margin,
                               mode=mode)
        assert torch.equal(r.status, alone.status), k
        assert torch.equal(r.steps, alone.steps), k
        assert torch.equal(r.attempts, alone.attempts), k
        assert torch.equal(r.states.view(torch.int64), alone.states.view(torch.int64)), k
        assert r.iterations == alone.iterations, k
        assert torch.equal(r.states[:, 4].view(torch.int64), tj.state0[:, 4].view(torch.int64))
        # the reference's trace_rays on 12 seeded rays of every job, with
        # the host-to-host spread of each stored beside it
        idx = torch.from_numpy(g[f"j{k}_index"]).cuda()
        np.testing.assert_array_equal(tj.state0[idx].cpu().numpy(), g[f"j{k}_state0"])
        assert_trace_parity(r.states[idx].cpu().numpy(), r.status[idx].cpu().numpy(),
                            r.steps[idx].cpu().numpy(), g[f"j{k}_states"],
                            g[f"j{k}_status"], g[f"j{k}_steps"],
                            label=f"uneven job {k} in the full launch ({mode})",
                            spread=g[f"j{k}_spread"])
