"""Pins the oracle's trace_rays to the reference: golden vectors written by
the reference's own trace_rays (tests/golden/make_golden.py), the shadow
boundary closed forms of the reference's test_shadow.py, and the capture
mask of the reference-held image shadow_gpu.png."""

import math

import numpy as np
import pytest

from conftest import golden
from oracle import bh_oracle as orc
from parity import assert_trace_parity
from blackhorizon_amd import workloads

EDGE_CASES = ["default", "nonfinite", "max_steps_1", "max_steps_7", "min_step_high", "loose",
              "tight_small_h", "max_steps_0"]
EDGE_CASES_R8 = ["default", "nonfinite", "nonfinite_pos", "max_steps_0", "max_steps_neg",
                 "max_steps_1", "max_steps_7", "attempt_cap", "min_step_high", "small_h", "huge_h"]


def _kw(g, name):
    pre = name + "_kw_"
    kw = {k[len(pre):]: float(g[k]) for k in g.files if k.startswith(pre)}
    if "max_steps" in kw:
        kw["max_steps"] = int(kw["max_steps"])
    return kw


@pytest.mark.parametrize("case", EDGE_CASES)
def test_edge_cases_match_reference(case):
    g = golden("edge_cases.npz")
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        res = orc.trace_rays(orc.Kerr(spin=0.9), g[case + "_state0"], **_kw(g, case))
    assert res.iterations == int(g[case + "_iterations"])
    assert_trace_parity(res.states, res.status, res.steps, g[case + "_states"],
                        g[case + "_status"], g[case + "_steps"], label=case)


@pytest.mark.parametrize("case", EDGE_CASES_R8)
def test_edge_cases_r8_match_reference(case):
    """The same outcomes at rtol 1e-8 (edge_cases_r8.npz), bit for bit."""
    g = golden("edge_cases_r8.npz")
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        res = orc.trace_rays(orc.Kerr(spin=0.9), g[case + "_state0"], **_kw(g, case))
    assert res.iterations == int(g[case + "_iterations"])
    np.testing.assert_array_equal(res.status, g[case + "_status"])
    np.testing.assert_array_equal(res.steps, g[case + "_steps"])
    np.testing.assert_array_equal(res.states, g[case + "_states"])


def test_small_frame_subsample_matches_reference():
    g = golden("small_frame_stride8.npz")
    job = workloads.small_frame()
    res = orc.trace_rays(orc.Kerr(spin=job.spin), g["state0"], job.escape_radius,
                         max_steps=job.max_steps, rtol=job.rtol, atol=job.atol,
                         capture_margin=job.capture_margin)
    assert res.iterations == int(g["iterations"])
    assert_trace_parity(res.states, res.status, res.steps, g["states"], g["status"], g["steps"],
                        label="small frame")


def test_fans_match_reference():
    g = golden("fans_d30.npz")
    for spin in (0.9, 0.998):
        k = f"a{spin}"
        res = orc.trace_rays(orc.Kerr(spin=spin), g[k + "_state0"], 60.0, rtol=1e-10)
        assert_trace_parity(res.states, res.status, res.steps, g[k + "_states"],
                            g[k + "_status"], g[k + "_steps"], label=k)


def test_uneven_mix_subsample_matches_reference():
    g = golden("uneven_subsample.npz")
    for j, job in enumerate(workloads.uneven_mix()[:8]):
        res = orc.trace_rays(orc.Kerr(spin=job.spin), g[f"j{j}_state0"], job.escape_radius,
                             max_steps=job.max_steps, rtol=job.rtol, atol=job.atol,
                             capture_margin=job.capture_margin)
        assert res.iterations == int(g[f"j{j}_iterations"])
        assert_trace_parity(res.states, res.status, res.steps, g[f"j{j}_states"],
                            g[f"j{j}_status"], g[f"j{j}_steps"], label=f"job {j}")


def test_uneven_all_jobs_match_reference():
    """12 rays of each of the 64 uneven-mix jobs (uneven_all.npz): the oracle
    reproduces the reference's statuses, steps and final states, every ray
    within 10x its stored host-to-host spread where that exceeds 1e-7, and
    bit for bit on every ray whose spread is zero."""
    g = golden("uneven_all.npz")
    for j, job in enumerate(workloads.uneven_mix()):
        res = orc.trace_rays(orc.Kerr(spin=job.spin), g[f"j{j}_state0"], job.escape_radius,
                             max_steps=job.max_steps, rtol=job.rtol, atol=job.atol,
                             capture_margin=job.capture_margin)
        assert_trace_parity(res.states, res.status, res.steps, g[f"j{j}_states"],
                            g[f"j{j}_status"], g[f"j{j}_steps"], label=f"job {j}",
                            spread=g[f"j{j}_spread"])
        exact = g[f"j{j}_spread"] == 0.0
        np.testing.assert_array_equal(res.states[exact], g[f"j{j}_states"][exact])


# --- closed forms of the reference's test_shadow.py -------------------------

def fan(st, offsets, distance=1000.0):
    n = offsets.shape[0]
    pos = np.tile([[distance, 0.0, 0.0]], (n, 1))
    d = np.stack([-np.ones(n), offsets, np.zeros(n)], axis=-1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, d, "past"))


def capture_boundary(trace, st, lo, hi, rounds=3, n=48):
    """Bisect |b| = |L_z / E| at the capture/escape edge of an equatorial fan."""
    b = None
    for _ in range(rounds):
        off = np.linspace(lo, hi, n)
        s0 = fan(st, off)
        status = np.asarray(trace(s0))
        e, lz = orc.conserved_quantities(s0)
        imp = lz / e
        cap = status == 1
        esc = status == 2
        assert cap.any() and esc.any() and (cap | esc).all()
        mag = np.abs(off)
        ic = int(np.argmax(np.where(cap, mag, -1.0)))
        lo = float(off[ic])
        ie = int(np.argmin(np.where(esc & (mag > mag[ic]), mag, np.inf)))
        hi = float(off[ie])
        b = 0.5 * (abs(imp[ic]) + abs(imp[ie]))
    return b


def bardeen_b(spin, prograde):
    st = orc.Kerr(spin=spin)
    r = st.photon_orbit_radius(prograde)
    return abs((r * r * (r - 3.0) + spin * spin * (r + 1.0)) / (spin * (1.0 - r)))


def oracle_fan_trace(spin):
    st = orc.Kerr(spin=spin)
    return st, lambda s0: orc.trace_rays(st, s0, 1200.0, max_steps=60000, rtol=1e-10,
                                         atol=1e-13, max_step=5.0).status


def test_oracle_schwarzschild_shadow():
    st, tr = oracle_fan_trace(0.0)
    b = capture_boundary(tr, st, 4.5e-3, 6.5e-3)
    assert abs(b - 3.0 * math.sqrt(3.0)) / (3.0 * math.sqrt(3.0)) < 5e-3


def test_oracle_kerr_shadow_asymmetry():
    st, tr = oracle_fan_trace(0.9)
    bp = capture_boundary(tr, st, -4.0e-3, -2.0e-3)
    br = capture_boundary(tr, st, 5.5e-3, 8.0e-3)
    assert abs(bp - bardeen_b(0.9, True)) / bardeen_b(0.9, True) < 5e-3
    assert abs(br - bardeen_b(0.9, False)) / bardeen_b(0.9, False) < 5e-3
    assert bp < br


def test_oracle_shadow_png_subsample():
    """Capture/escape and checker class of every 11th pixel of shadow_gpu.png."""
    cls = golden("shadow_gpu_classes.npz")["classes"].ravel()
    job = workloads.shadow_cli_frame()
    st = orc.Kerr(spin=job.spin)
    cam = orc.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees, width=job.width,
                                       height=job.height)
    s0 = orc.camera_state(st, cam)[::11]
    res = orc.trace_rays(st, s0, job.escape_radius, max_steps=job.max_steps, rtol=job.rtol)
    want = cls[::11]
    captured = res.status == 1
    np.testing.assert_array_equal(captured, want == 0)
    esc = res.status == 2
    assert np.all(esc == (want != 0))
    light = orc.checker_parity(st, res.states[esc])
    agree = np.mean(light == (want[esc] == 2))
    assert agree > 0.999, agree


def test_reference_spread_small_on_escaped_rays():
    """The host-to-host spread of the reference (NumPy SIMD pow vs libm pow)
    is tiny on escaped rays and only exceeds 1e-7 on captured rays, which end
    inside the horizon with momenta of order 1e3."""
    from parity import reference_spread

    g = golden("small_frame_stride8.npz")
    job = workloads.small_frame()
    spread = reference_spread(job.spin, g["state0"], g["states"],
                              escape_radius=job.escape_radius, rtol=job.rtol, atol=job.atol,
                              capture_margin=job.capture_margin)
    esc = g["status"] == 2
    assert spread[esc].max() < 1e-8
    assert spread.max() < 1e-6
    assert np.linalg.norm(g["states"][~esc, 5:8], axis=-1).min() > 50.0


def test_camera_tile_order_is_permutation():
    from blackhorizon_amd.camera import PinholeCamera

    for w, h, t in ((160, 120, 8), (13, 7, 4), (1, 1, 8), (64, 3, 16)):
        cam = PinholeCamera.from_orbit(30.0, 60.0, width=w, height=h)
        o = cam.tile_order(t)
        assert np.array_equal(np.sort(o), np.arange(w * h))
        if w == 160:
            first = o[:64]
            assert np.array_equal(np.unique(first // 160), np.arange(8))
            assert np.array_equal(np.unique(first % 160), np.arange(8))
        c = cam.center_tile_order(t)
        assert np.array_equal(np.sort(c), np.arange(w * h))
        if w == 160:
            # the first tile is one of the four around the image centre; the
            # last ones are corner tiles
            first = c[:64]
            assert set(np.unique(first // 160)) <= set(range(52, 68))
            assert set(np.unique(first % 160)) <= set(range(72, 88))
            last = c[-1]
            assert last // 160 in range(0, 8) or last // 160 in range(112, 120)
