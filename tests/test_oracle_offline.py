"""The offline tile tracer oracle (oracle/offline_oracle.py) against what
pins it without the reference: Schwarzschild disk hits against an
independent quadrature of the photon orbit equation, the tile loop without
a disk against the (separately pinned) tracer restatement, radial flat
directions, and the camera sample grid's geometry."""

import math

import numpy as np
import pytest
from scipy import integrate, optimize

from oracle import bh_oracle as orc
from oracle import offline_oracle as off


def _face_on_rays(D, alphas):
    """Photons leaving (0, 0, D) in the x-z plane at angle alpha from -z."""
    st = orc.Kerr(spin=0.0)
    d = np.stack([np.sin(alphas), np.zeros_like(alphas), -np.cos(alphas)], -1)
    pos = np.tile([0.0, 0.0, D], (len(alphas), 1))
    return st, orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, d))


def _quadrature_hit_radius(b, D):
    """r where a Schwarzschild photon of impact parameter b, starting at
    r = D on the axis and moving inward, has swept pi/2: the solution of
    int_r^D dr / (r^2 sqrt(1/b^2 - (1 - 2/r) / r^2)) = pi/2 (no turning
    point before the crossing)."""
    def swept(r_lo):
        f = lambda r: 1.0 / (r * r * math.sqrt(1.0 / (b * b) - (1.0 - 2.0 / r) / (r * r)))
        return integrate.quad(f, r_lo, D, epsabs=1e-13, epsrel=1e-12, limit=400)[0]

    # periapsis: largest root of r^3 - b^2 r + 2 b^2 = 0
    roots = np.roots([1.0, 0.0, -b * b, 2.0 * b * b])
    rp = max(r.real for r in roots if abs(r.imag) < 1e-12)
    return optimize.brentq(lambda r: swept(r) - math.pi / 2.0, rp * (1 + 1e-9), D * 0.999,
                           xtol=1e-14, rtol=1e-14)


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
def test_schwarzschild_face_on_disk_hits_match_quadrature():
    D = 30.0
    alphas = np.radians(np.linspace(11.0, 18.0, 8))  # crossings before periapsis
    st, s0 = _face_on_rays(D, alphas)
    out = off.trace_tile(st, s0, escape_radius=60.0, disk_radii=(2.5, 40.0), rtol=1e-10,
                         max_steps=20000, max_step=0.05)
    assert np.all(out.status == int(orc.RayStatus.DISK))
    b = D * np.abs(s0[:, 5])  # |L| / |E| with L = D p_x, |p_t| = 1
    for k in range(len(alphas)):
        r_q = _quadrature_hit_radius(b[k], D)
        r_hit = np.linalg.norm(out.hit_positions[k])
        assert abs(out.hit_positions[k, 2]) < 1e-9
        assert abs(r_hit - r_q) <= 1e-7 * r_q, (k, r_hit, r_q)


def test_disk_annulus_misses_continue():
    """A crossing outside [inner, outer] does not stop the ray."""
    D = 30.0
    st, s0 = _face_on_rays(D, np.radians([12.0, 14.0]))
    hit = off.trace_tile(st, s0, 60.0, (2.5, 40.0), 1e-9, 20000)
    r = np.linalg.norm(hit.hit_positions, axis=-1)
    miss = off.trace_tile(st, s0, 60.0, (r.max() * 1.01, 40.0), 1e-9, 20000)
    assert np.all(hit.status == 5)
    assert np.all(miss.status != 5)
    assert np.all(miss.hit_positions == 0.0)


def test_tile_loop_without_disk_equals_the_tracer():
    """Classify-then-step and the tracer's step-then-classify reach the same
    final state, steps and status on escaped and captured rays, given the
    tile loop's fixed h0 = 0.1 and atol = rtol 1e-3."""
    st = orc.Kerr(spin=0.9)
    cam = orc.PinholeCamera.from_orbit(30.0, 60.0, fov_degrees=30.0, width=40, height=30)
    s0 = orc.camera_state(st, cam)[::7]
    rtol = 1e-8
    tile = off.trace_tile(st, s0, 45.0, None, rtol, 20000, capture_margin=0.01 / 1.4358898944)
    ref = orc.trace_rays(st, s0, 45.0, max_steps=20000, rtol=rtol, atol=rtol * 1e-3,
                         initial_step=0.1, min_step=0.0, capture_margin=0.01 / 1.4358898944)
    assert set(np.unique(tile.status)) <= {1, 2}
    np.testing.assert_array_equal(tile.status, ref.status)
    np.testing.assert_array_equal(tile.steps, ref.steps)
    esc = ref.status == 2
    d = off.flat_directions(st, ref.states[esc][:, 1:4], ref.states[esc][:, 5:8],
                            p_t=ref.states[esc][:, 4])
    np.testing.assert_array_equal(tile.escape_directions[esc], d)
    assert np.all(tile.escape_directions[~esc] == 0.0)


def test_flat_direction_of_radial_rays_is_radial():
    st = orc.Kerr(spin=0.0)
    rng = np.random.default_rng(3)
    u = rng.normal(size=(20, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    pos = 80.0 * u
    for sign in (1.0, -1.0):
        s = orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, sign * u))
        d = off.flat_directions(st, s[:, 1:4], s[:, 5:8], p_t=s[:, 4])
        np.testing.assert_allclose(np.abs(np.sum(d * u, axis=-1)), 1.0, atol=1e-14)


def test_saturation_and_retrace_bookkeeping():
    st = orc.Kerr(spin=0.9)
    cam = off.FlyCamera.from_orbit(26.0, 82.0)
    dirs = off.subpixel_directions(cam, 16, 10, 70.0, 1)
    pos = np.tile(cam.position, (len(dirs), 1))
    s0 = orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, dirs))
    kw = dict(escape_radius=1.3 * 26.0, disk_radii=(st.isco_radius(True), 18.0))
    first = off.trace_tile(st, s0, rtol=1e-8, max_steps=60, **kw)
    sat = first.saturated
    assert np.array_equal(sat, (first.status == 3) | (first.steps >= 51))
    res, mask = off.trace_tile_retrace(st, s0, rtol=1e-8, refine_rtol=1e-10, max_steps=60, **kw)
    assert np.array_equal(mask, sat & (first.status != 5) & (first.status != 6))
    assert mask.any() and (~mask).any()
    alone = off.trace_tile(st, s0[mask], rtol=1e-10, max_steps=240, **kw)
    np.testing.assert_array_equal(res.status[mask], alone.status)
    np.testing.assert_array_equal(res.status[~mask], first.status[~mask])
    np.testing.assert_array_equal(res.saturated, first.saturated)


def test_subpixel_grid_geometry():
    cam = off.FlyCamera.from_orbit(25.0, 80.0)
    d = off.subpixel_directions(cam, 8, 6, 70.0, 2)
    assert d.shape == (8 * 6 * 4, 3)
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), 1.0, atol=1e-12)
    fwd = -cam.position / np.linalg.norm(cam.position)
    # the four central subpixels straddle the optical axis symmetrically
    grid = off.pinhole_grid(8, 6, 70.0, 2).reshape(6, 8, 2, 2, 3)
    np.testing.assert_allclose(grid[:, :, :, :, 1].sum(), 0.0, atol=1e-12)
    np.testing.assert_allclose(grid[:, :, :, :, 2].sum(), 0.0, atol=1e-12)
    # the field of view spans the width (render.py:136-142 scales u by
    # tan(fov / 2) and v by tan(fov / 2) / aspect): the outermost column's
    # samples sit just inside 35 deg, the top row's at 35 deg / aspect
    right = grid[:, -1, :, -1, :].reshape(-1, 3)
    az = np.degrees(np.arctan2(right[:, 1], right[:, 0]))
    expect = np.degrees(np.arctan((1.0 - 0.25 / 8 * 2) * math.tan(math.radians(35.0))))
    np.testing.assert_allclose(az, expect, rtol=1e-12)
    top = grid[0, :, 0, :, :].reshape(-1, 3)
    tan_el = top[:, 2] / top[:, 0]
    np.testing.assert_allclose(tan_el, (1.0 - 0.25 / 6 * 2) * math.tan(math.radians(35.0))
                               / (8 / 6), rtol=1e-12)
    assert np.dot(cam.basis()[0], fwd) > 1.0 - 1e-12


# ---------------------------------------------------------------------------
# the real-time shader's algorithm: the reference's own accuracy checks
# (test_realtime_reference.py:56-129) applied to the oracle
# ---------------------------------------------------------------------------

def _black(status):
    return (status == 1) | (status == 3)


def test_shader_presets_match_adaptive_tracer_for_kerr():
    st = orc.Kerr(spin=0.9)
    offsets = np.linspace(0.03, 0.40, 300)
    s0, _ = off.equatorial_fan(st, np.concatenate([offsets, -offsets]))
    truth = orc.trace_rays(st, s0, escape_radius=60.0, rtol=1e-10)
    for name, (max_steps, scale) in off.PRESETS.items():
        ref = off.trace_like_shader(st, s0, 60.0, max_steps=max_steps, step_scale=scale)
        assert np.array_equal(_black(ref.status), _black(truth.status)), name


def test_shader_schwarzschild_boundary_medium():
    st = orc.Kerr(spin=0.0)
    s0, impact = off.equatorial_fan(st, np.linspace(0.10, 0.25, 600))
    ref = off.trace_like_shader(st, s0, 60.0, *off.PRESETS["medium"])
    black = _black(ref.status)
    b = 0.5 * (np.abs(impact)[black].max() + np.abs(impact)[~black].min())
    assert abs(b - 3.0 * np.sqrt(3.0)) / (3.0 * np.sqrt(3.0)) < 5e-3


def test_shader_near_extremal_high_preset():
    """test_realtime_reference.py:90-108 on the oracles: at a = 0.998 the
    HIGH preset shades an equatorial fan like the adaptive tracer at rtol
    1e-10, with both classes on each side of the hole."""
    st = orc.Kerr(spin=0.998)
    offsets = np.linspace(0.03, 0.40, 300)
    s0, _ = off.equatorial_fan(st, np.concatenate([offsets, -offsets]))
    truth = orc.trace_rays(st, s0, 60.0, rtol=1e-10)
    ref = off.trace_like_shader(st, s0, 60.0, *off.PRESETS["high"])
    assert np.array_equal(_black(ref.status), _black(truth.status))
    for side in (slice(0, 300), slice(300, 600)):
        assert 0 < int(_black(ref.status[side]).sum()) < 300


def test_shader_plungers_never_fail_numerically():
    st = orc.Kerr(spin=0.998)
    s0, _ = off.equatorial_fan(st, np.linspace(0.001, 0.15, 200))
    ref = off.trace_like_shader(st, s0, 60.0, *off.PRESETS["low"])
    assert np.all(np.isfinite(ref.positions))
    assert np.all(ref.status == 1)


def test_shader_disk_interpolation_lies_on_the_chord():
    """Disk hits are the linear interpolation of the crossing step: z = 0 at
    the hit and the hit radius inside the annulus."""
    st = orc.Kerr(spin=0.9)
    cam = off.FlyCamera.from_orbit(26.0, 80.0)
    dirs = off.subpixel_directions(cam, 24, 16, 60.0, 1)
    pos = np.tile(cam.position, (len(dirs), 1))
    s0 = orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, dirs))
    res = off.trace_like_shader(st, s0, 60.0, disk_radii=(st.isco_radius(True), 18.0))
    hit = res.status == 5
    assert hit.sum() > 20
    assert np.max(np.abs(res.hit_positions[hit, 2])) < 1e-12
    assert np.all((res.hit_radii[hit] >= st.isco_radius(True)) & (res.hit_radii[hit] <= 18.0))
    assert np.all(np.isnan(res.hit_radii[~hit]))
