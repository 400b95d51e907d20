"""Write the golden trace fixtures under tests/golden/.

Run in a container that holds the reference checkout:

    python tests/golden/make_golden.py /root/reference [part ...]

Only the reference's ``trace_rays`` (tracer.py:50-157) and
``integrate_dense`` (integrators.py:127-174) are executed: they are the
functions on the path that no closed form or invariant pins to the per-ray
level (final states, status, accepted steps, iterations). Its input
states are built by the oracle (camera + null momenta), so the fixtures hold
plain data: ``state0`` rows and the reference outputs for them. Rays are
independent in ``trace_rays``, so a subsample of a frame has the same per-ray
results as the full frame; ``iterations`` is the subsample batch's value.

The capture/escape/checker class of every pixel of the reference-held image
``shadow_gpu.png`` is decoded from its pixel colours (imaging.py:20-24) into
``shadow_gpu_classes.npz``; that needs only the image file.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import bh_oracle as orc  # noqa: E402
from blackhorizon_amd import workloads  # noqa: E402


def _job_state(job):
    st = orc.Kerr(job.mass, job.spin)
    cam = orc.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees,
                                       width=job.width, height=job.height)
    return st, orc.camera_state(st, cam)


def _trace_ref(ref_trace, ref_kerr, job, state0, **over):
    kw = dict(escape_radius=job.escape_radius, max_steps=job.max_steps,
              rtol=job.rtol, atol=job.atol, initial_step=job.initial_step,
              max_step=job.max_step, min_step=job.min_step,
              capture_margin=job.capture_margin)
    kw.update(over)
    res = ref_trace(ref_kerr(mass=job.mass, spin=job.spin), state0, **kw)
    return res, kw


def frame_subsample(ref_trace, ref_kerr, job, stride, name):
    _, s0 = _job_state(job)
    idx = np.arange(0, s0.shape[0], stride)
    state0 = s0[idx].copy()
    res, kw = _trace_ref(ref_trace, ref_kerr, job, state0)
    out = dict(index=idx, state0=state0, states=np.asarray(res.states),
               status=np.asarray(res.status), steps=np.asarray(res.steps),
               iterations=np.int64(res.iterations), spin=job.spin, mass=job.mass,
               **{k: np.float64(v) for k, v in kw.items() if k != "max_steps"},
               max_steps=np.int64(kw["max_steps"]))
    np.savez_compressed(os.path.join(HERE, name), **out)
    print(name, state0.shape[0], "rays, iterations", res.iterations,
          "status counts", np.bincount(res.status, minlength=7))


def fans(ref_trace, ref_kerr):
    """Equatorial fans of tests/test_realtime_reference.py:161-174 (D = 30)."""
    rows = {}
    for spin in (0.9, 0.998):
        st = orc.Kerr(spin=spin)
        off = np.linspace(0.03, 0.40, 300)
        off = np.concatenate([off, -off])
        n = off.shape[0]
        pos = np.tile([[30.0, 0.0, 0.0]], (n, 1))
        d = np.stack([-np.ones(n), off, np.zeros(n)], axis=-1)
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        state0 = orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, d, "past"))
        res = ref_trace(ref_kerr(spin=spin), state0, escape_radius=60.0, rtol=1e-10)
        key = f"a{spin}"
        rows[key + "_state0"] = state0
        rows[key + "_states"] = np.asarray(res.states)
        rows[key + "_status"] = np.asarray(res.status)
        rows[key + "_steps"] = np.asarray(res.steps)
        rows[key + "_iterations"] = np.int64(res.iterations)
    np.savez_compressed(os.path.join(HERE, "fans_d30.npz"), **rows)
    print("fans_d30.npz")


def edge_cases(ref_trace, ref_kerr):
    """Host-forceable outcomes of the controller (tracer.py:122-154)."""
    st = orc.Kerr(spin=0.9)
    pos = np.array([[20.0, 1.0, 2.0], [20.0, 0.0, 0.0], [10.0, 3.0, -4.0],
                    [1.2, 0.1, 0.3], [20.0, 0.5, 0.5], [8.0, -2.0, 1.0]])
    d = np.array([[-1.0, 0.1, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.2],
                  [-1.0, 0.0, 0.0], [-1.0, 0.0, 0.05], [0.3, 1.0, -0.1]])
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    # the ray inside the horizon (row 3) gets its momentum from a point outside
    mom = orc.null_momentum_from_velocity(st, np.where(np.arange(6)[:, None] == 3,
                                                       np.array([[5.0, 0.1, 0.3]]), pos), d, "past")
    base = orc.build_state(pos, mom)
    bad = base.copy()
    bad[0, 5] = np.nan   # momentum NaN
    bad[2, 6] = np.inf   # momentum inf (an infinite position would fail the r0 check)
    bad[4, 0] = np.nan   # coordinate time NaN: RHS finite, state not
    cases = {
        "default": (base, dict(escape_radius=40.0)),
        "nonfinite": (bad, dict(escape_radius=40.0)),
        "max_steps_1": (base, dict(escape_radius=40.0, max_steps=1)),
        "max_steps_7": (base, dict(escape_radius=40.0, max_steps=7, rtol=1e-12, atol=1e-14)),
        "min_step_high": (base, dict(escape_radius=40.0, min_step=0.5)),
        "loose": (base, dict(escape_radius=40.0, rtol=1e-4, atol=1e-6, max_step=5.0)),
        "tight_small_h": (base, dict(escape_radius=40.0, rtol=1e-12, atol=1e-14,
                                     initial_step=1e-3, max_step=0.5)),
        "max_steps_0": (base, dict(escape_radius=40.0, max_steps=0)),
    }
    out = {}
    for name, (s0, kw) in cases.items():
        res = ref_trace(ref_kerr(spin=0.9), s0, **kw)
        out[name + "_state0"] = s0
        out[name + "_states"] = np.asarray(res.states)
        out[name + "_status"] = np.asarray(res.status)
        out[name + "_steps"] = np.asarray(res.steps)
        out[name + "_iterations"] = np.int64(res.iterations)
        for k, v in kw.items():
            out[name + "_kw_" + k] = np.float64(v)
        print(name, res.status, res.steps, res.iterations)
    np.savez_compressed(os.path.join(HERE, "edge_cases.npz"), **out)


def edge_cases_r8(ref_trace, ref_kerr):
    """The controller's failure and cap outcomes (tracer.py:101-154) at
    rtol 1e-8, atol 1e-10, where the fast build runs: non-finite momenta,
    time and positions (FAILED), max_steps 0, negative, 1 and 7 (MAX_STEPS
    by accepted steps), an initial step so large that some rays hit the
    4 max_steps attempt cap first, h below min_step (FAILED), a small
    initial and maximum step, and a huge one."""
    st = orc.Kerr(spin=0.9)
    pos = np.array([[20.0, 1.0, 2.0], [20.0, 0.0, 0.0], [10.0, 3.0, -4.0],
                    [1.2, 0.1, 0.3], [20.0, 0.5, 0.5], [8.0, -2.0, 1.0]])
    d = np.array([[-1.0, 0.1, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.2],
                  [-1.0, 0.0, 0.0], [-1.0, 0.0, 0.05], [0.3, 1.0, -0.1]])
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    mom = orc.null_momentum_from_velocity(st, np.where(np.arange(6)[:, None] == 3,
                                                       np.array([[5.0, 0.1, 0.3]]), pos), d, "past")
    base = orc.build_state(pos, mom)
    bad = base.copy()
    bad[0, 5] = np.nan   # momentum NaN
    bad[2, 6] = np.inf   # momentum inf
    bad[4, 0] = np.nan   # coordinate time NaN: RHS finite, state not
    bad_pos = base.copy()
    bad_pos[1, 1] = np.nan    # position NaN: r0 NaN passes the r0 check
    bad_pos[3, 7] = -np.inf   # momentum -inf on the ray inside the horizon
    bad_pos[5, 4] = np.nan    # p_t NaN
    r8 = dict(escape_radius=40.0, rtol=1e-8, atol=1e-10)
    cases = {
        "default": (base, {}),
        "nonfinite": (bad, {}),
        "nonfinite_pos": (bad_pos, {}),
        "max_steps_0": (base, dict(max_steps=0)),
        "max_steps_neg": (base, dict(max_steps=-3)),
        "max_steps_1": (base, dict(max_steps=1)),
        "max_steps_7": (base, dict(max_steps=7)),
        "attempt_cap": (base, dict(max_steps=2, initial_step=1e4, max_step=1e4)),
        "min_step_high": (base, dict(min_step=0.5)),
        "small_h": (base, dict(initial_step=1e-3, max_step=0.5)),
        "huge_h": (base, dict(initial_step=1e3, max_step=1e3, max_steps=50)),
    }
    out = {}
    for name, (s0, over) in cases.items():
        kw = dict(r8)
        kw.update(over)
        res = ref_trace(ref_kerr(spin=0.9), s0, **kw)
        out[name + "_state0"] = s0
        out[name + "_states"] = np.asarray(res.states)
        out[name + "_status"] = np.asarray(res.status)
        out[name + "_steps"] = np.asarray(res.steps)
        out[name + "_iterations"] = np.int64(res.iterations)
        for k, v in kw.items():
            out[name + "_kw_" + k] = np.float64(v)
        print(name, res.status, res.steps, res.iterations)
    np.savez_compressed(os.path.join(HERE, "edge_cases_r8.npz"), **out)


def uneven_subsample(ref_trace, ref_kerr, n_jobs=8, per_job=48):
    """A few rays from each of the first jobs of the uneven mix (mixed
    spins of both signs, distances and tolerances)."""
    out = {}
    for j, job in enumerate(workloads.uneven_mix()[:n_jobs]):
        _, s0 = _job_state(job)
        rng = np.random.default_rng(1000 + j)
        idx = np.sort(rng.choice(s0.shape[0], size=min(per_job, s0.shape[0]), replace=False))
        state0 = s0[idx].copy()
        res, kw = _trace_ref(ref_trace, ref_kerr, job, state0)
        out[f"j{j}_index"] = idx
        out[f"j{j}_state0"] = state0
        out[f"j{j}_states"] = np.asarray(res.states)
        out[f"j{j}_status"] = np.asarray(res.status)
        out[f"j{j}_steps"] = np.asarray(res.steps)
        out[f"j{j}_iterations"] = np.int64(res.iterations)
        print("uneven job", j, job.spin, job.distance, job.rtol, np.bincount(res.status, minlength=7))
    np.savez_compressed(os.path.join(HERE, "uneven_subsample.npz"), **out)


def uneven_all(ref_trace, ref_kerr, per_job=12):
    """Seeded rays from every job of the 64-job uneven mix, traced by the
    reference's trace_rays with each job's settings, plus ``spread``: per
    ray, how far the final state moves under ulp-level step-factor changes
    (the oracle with libm pow and with random-ulp pow), precomputed here so
    the full-size GPU test need not re-trace on the host."""
    import warnings

    out = {}
    for j, job in enumerate(workloads.uneven_mix()):
        _, s0 = _job_state(job)
        rng = np.random.default_rng(3000 + j)
        idx = np.sort(rng.choice(s0.shape[0], size=min(per_job, s0.shape[0]), replace=False))
        state0 = s0[idx].copy()
        res, kw = _trace_ref(ref_trace, ref_kerr, job, state0)
        states = np.asarray(res.states)
        spread = np.zeros(len(idx))
        st = orc.Kerr(job.mass, job.spin)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            for fn in (orc.step_factor_libm, orc.step_factor_ulp_noise(1),
                       orc.step_factor_ulp_noise(2)):
                alt = orc.trace_rays(st, state0, factor_fn=fn, **kw)
                for b in (slice(0, 1), slice(1, 4), slice(4, 5), slice(5, 8)):
                    d = np.linalg.norm(alt.states[:, b] - states[:, b], axis=-1) / np.maximum(
                        np.linalg.norm(states[:, b], axis=-1), 1.0)
                    spread = np.maximum(spread, np.nan_to_num(d, nan=np.inf))
        out[f"j{j}_index"] = idx
        out[f"j{j}_state0"] = state0
        out[f"j{j}_states"] = states
        out[f"j{j}_status"] = np.asarray(res.status)
        out[f"j{j}_steps"] = np.asarray(res.steps)
        out[f"j{j}_spread"] = spread
        print("uneven job", j, job.rtol, np.bincount(res.status, minlength=7), "max spread",
              spread.max(), flush=True)
    np.savez_compressed(os.path.join(HERE, "uneven_all.npz"), **out)


def hd_subsample(ref_trace, ref_kerr, per_frame=96):
    """Seeded rays of each 1920x1080 frame (spins x inclinations)."""
    out = {}
    for f, job in enumerate(workloads.hd_frames()):
        st = orc.Kerr(job.mass, job.spin)
        cam = orc.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                           fov_degrees=job.fov_degrees,
                                           width=job.width, height=job.height)
        rng = np.random.default_rng(2000 + f)
        idx = np.sort(rng.choice(job.n_rays, size=per_frame, replace=False))
        pos = cam.ray_origins()[idx]
        dirs = cam.ray_directions()[idx]
        state0 = orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, dirs, "past"))
        res, kw = _trace_ref(ref_trace, ref_kerr, job, state0)
        out[f"f{f}_index"] = idx
        out[f"f{f}_state0"] = state0
        out[f"f{f}_states"] = np.asarray(res.states)
        out[f"f{f}_status"] = np.asarray(res.status)
        out[f"f{f}_steps"] = np.asarray(res.steps)
        print("hd frame", f, job.spin, job.inclination_degrees, np.bincount(res.status, minlength=7))
    np.savez_compressed(os.path.join(HERE, "hd_subsample.npz"), **out)


def shadow_png_classes(ref_root):
    """0 = shadow, 1 = dark checker, 2 = light checker, 3 = other colour."""
    from PIL import Image

    img = np.asarray(Image.open(os.path.join(ref_root, "shadow_gpu.png")).convert("RGB"))
    cls = np.full(img.shape[:2], 3, dtype=np.uint8)
    cls[np.all(img == [0, 0, 0], axis=-1)] = 0
    cls[np.all(img == [24, 30, 52], axis=-1)] = 1
    cls[np.all(img == [104, 128, 178], axis=-1)] = 2
    np.savez_compressed(os.path.join(HERE, "shadow_gpu_classes.npz"), classes=cls)
    print("shadow_gpu_classes", np.bincount(cls.ravel(), minlength=4))


def dense_orbits(integrate_dense, geodesic_rhs, ref_kerr):
    """integrate_dense (integrators.py:127-174) run on one ray at a time,
    where its batch-shared step is that ray's own: the per-ray reference for
    integrating to a fixed affine parameter.  Rays: the photon flyby and the
    circular r = 6 M orbit of the reference's test_orbits.py (22-46,
    108-136) and a few seeded near-circular orbits around a = 0.9."""
    st = orc.Kerr(spin=0.9)
    cases = []
    pos = np.array([[30.0, 4.0, 1.5]])
    mom = orc.null_momentum_from_velocity(st, pos, np.array([[-1.0, 0.0, 0.0]]), "future")
    cases.append((orc.build_state(pos, mom), 70.0, 1e-11))
    om = 1.0 / (6.0 ** 1.5 + 0.9)
    pos = np.array([[6.0, 0.9, 0.0]])
    mom = orc.timelike_momentum(st, pos, np.array([[-om * 0.9, om * 6.0, 0.0]]))
    cases.append((orc.build_state(pos, mom), 170.0, 1e-11))
    p, v = workloads.near_circular_orbits(12, seed=11)
    for i in range(12):
        mom = orc.timelike_momentum(st, p[i:i + 1], v[i:i + 1])
        cases.append((orc.build_state(p[i:i + 1], mom), 400.0, 1e-10))
    out = {}
    kst = ref_kerr(spin=0.9)
    k = 0
    for s0, lam, rtol in cases:
        try:
            states = integrate_dense(lambda y: geodesic_rhs(kst, y), s0, lambda_total=lam,
                                     rtol=rtol)
        except RuntimeError as e:
            # the last step can leave a remainder of a few ulp of lambda, whose
            # follow-up step underflows; whether that happens is decided by
            # rounding, so such rays are not fixtures
            print("skipped:", e)
            continue
        out[f"c{k}_state0"] = s0
        out[f"c{k}_final"] = np.asarray(states[-1])
        out[f"c{k}_steps"] = np.int64(len(states) - 1)
        out[f"c{k}_lambda"] = np.float64(lam)
        out[f"c{k}_rtol"] = np.float64(rtol)
        print("dense orbit", k, len(states) - 1)
        k += 1
    np.savez_compressed(os.path.join(HERE, "dense_orbits.npz"), **out)


def c5_orbits(integrate_dense, geodesic_rhs, ref_kerr, n_pick=8):
    """Seeded orbits from inside the timelike-orbit workload (65,536
    near-circular orbits around a = 0.9 in 16 jobs, proper time 1e4 M,
    rtol 1e-10), each integrated alone by the reference's integrate_dense
    (integrators.py:127-174) with that workload's settings (its defaults:
    atol 1e-12, initial step 0.05, max step 2).  ``spread`` is how far the
    final state moves under ulp-level step-factor changes (the oracle with
    libm pow and with random-ulp pow), the conditioning of each orbit's
    result."""
    import warnings

    jobs = workloads.orbit_batches()
    pos = np.concatenate([j.positions for j in jobs])
    vel = np.concatenate([j.velocities for j in jobs])
    st = orc.Kerr(spin=0.9)
    kst = ref_kerr(spin=0.9)
    rng = np.random.default_rng(2024)
    out = {}
    k = 0
    for i in rng.permutation(len(pos)):
        if k == n_pick:
            break
        s0 = orc.build_state(pos[i:i + 1], orc.timelike_momentum(st, pos[i:i + 1],
                                                                 vel[i:i + 1]))
        try:
            states = integrate_dense(lambda y: geodesic_rhs(kst, y), s0, lambda_total=1.0e4,
                                     rtol=1e-10)
        except RuntimeError as e:
            print("skipped", i, e)
            continue
        final = np.asarray(states[-1])
        if np.linalg.norm(final[0, 1:4]) > 900.0 or np.linalg.norm(final[0, 1:4]) < 2.0:
            print("skipped", i, "(escapes or plunges)")
            continue
        spread = 0.0
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            for fn in (orc.step_factor_libm, orc.step_factor_ulp_noise(1),
                       orc.step_factor_ulp_noise(2)):
                alt = orc.integrate_dense(lambda y: orc.geodesic_rhs(st, y), s0, 1.0e4,
                                          rtol=1e-10, factor_fn=fn)[-1]
                d = np.max([np.linalg.norm(alt[0, b] - final[0, b]) /
                            max(np.linalg.norm(final[0, b]), 1.0)
                            for b in (slice(0, 1), slice(1, 4), slice(4, 5), slice(5, 8))])
                spread = max(spread, float(d))
        out[f"o{k}_index"] = np.int64(i)
        out[f"o{k}_state0"] = s0
        out[f"o{k}_final"] = final
        out[f"o{k}_steps"] = np.int64(len(states) - 1)
        out[f"o{k}_spread"] = np.float64(spread)
        print("c5 orbit", k, "index", i, "steps", len(states) - 1, "spread", spread, flush=True)
        k += 1
    np.savez_compressed(os.path.join(HERE, "c5_orbits.npz"), **out)


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    parts = sys.argv[2:] or ["png", "edge", "edge_r8", "fans", "small", "uneven", "uneven_all",
                             "hd", "dense", "c5"]
    sys.path.insert(0, ref_root)
    from blackhorizon.geodesics import geodesic_rhs
    from blackhorizon.integrators import integrate_dense
    from blackhorizon.kerr import KerrSpacetime
    from blackhorizon.tracer import trace_rays

    if "png" in parts:
        shadow_png_classes(ref_root)
    if "edge" in parts:
        edge_cases(trace_rays, KerrSpacetime)
    if "edge_r8" in parts:
        edge_cases_r8(trace_rays, KerrSpacetime)
    if "fans" in parts:
        fans(trace_rays, KerrSpacetime)
    if "small" in parts:
        frame_subsample(trace_rays, KerrSpacetime, workloads.small_frame(), 8,
                        "small_frame_stride8.npz")
    if "uneven" in parts:
        uneven_subsample(trace_rays, KerrSpacetime)
    if "hd" in parts:
        hd_subsample(trace_rays, KerrSpacetime)
    if "uneven_all" in parts:
        uneven_all(trace_rays, KerrSpacetime)
    if "dense" in parts:
        dense_orbits(integrate_dense, geodesic_rhs, KerrSpacetime)
    if "c5" in parts:
        c5_orbits(integrate_dense, geodesic_rhs, KerrSpacetime)


if __name__ == "__main__":
    main()
