"""GPU: the real-time shader's algorithm (bh_trace_shader_f64) bit for bit
against the oracle of realtime/reference.py, and the reference's own
accuracy checks against the adaptive tracer (test_realtime_reference.py)."""

import numpy as np
import pytest

from oracle import bh_oracle as orc
from oracle import offline_oracle as off

pytestmark = pytest.mark.gpu


def _settings(preset=None, **kw):
    from blackhorizon_amd.realtime import QualityPreset, RenderSettings

    s = RenderSettings(**kw)
    return s.apply_preset(QualityPreset(preset)) if preset else s


def _black(status):
    return (status == 1) | (status == 3)


@pytest.mark.parametrize("preset", ["low", "medium", "high", "ultra"])
def test_shader_bitwise_equals_oracle_on_fans(hip, preset):
    from blackhorizon_amd import realtime

    st = orc.Kerr(spin=0.9)
    offsets = np.linspace(0.03, 0.40, 150)
    s0, _ = off.equatorial_fan(st, np.concatenate([offsets, -offsets]))
    ref = off.trace_like_shader(st, s0, 60.0, *off.PRESETS[preset])
    got = realtime.trace_like_shader(hip.KerrSpacetime(spin=0.9), s0, _settings(preset), 60.0)
    np.testing.assert_array_equal(got.status, ref.status)
    np.testing.assert_array_equal(got.positions, ref.positions)
    np.testing.assert_array_equal(got.momenta, ref.momenta)


def test_shader_disk_and_interior_bitwise(hip):
    from blackhorizon_amd import realtime

    st = orc.Kerr(spin=0.9)
    cam = off.FlyCamera.from_orbit(26.0, 80.0)
    dirs = off.subpixel_directions(cam, 24, 16, 60.0, 1)
    pos = np.tile(cam.position, (len(dirs), 1))
    s0 = orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, dirs))
    disk = (st.isco_radius(True), 18.0)
    ref = off.trace_like_shader(st, s0, 60.0, disk_radii=disk)
    got = realtime.trace_like_shader(hip.KerrSpacetime(spin=0.9), s0, _settings(), 60.0,
                                     disk_radii=disk)
    assert (ref.status == 5).sum() > 20
    for name in ("status", "positions", "momenta", "hit_radii", "hit_positions", "hit_momenta"):
        np.testing.assert_array_equal(getattr(got, name), getattr(ref, name), err_msg=name)
    # interior camera: rays from r = 1.2 inside a = 0.5's horizon
    st = orc.Kerr(spin=0.5)
    rng = np.random.default_rng(9)
    u = rng.normal(size=(128, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    d = rng.normal(size=(128, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    p = 1.2 * u
    s1 = orc.build_state(p, orc.null_momentum_from_velocity(st, p, d, normalize=False))
    ref = off.trace_like_shader(st, s1, 60.0, interior_stop=0.02)
    got = realtime.trace_like_shader(hip.KerrSpacetime(spin=0.5), s1, _settings(), 60.0,
                                     interior_stop=0.02)
    np.testing.assert_array_equal(got.status, ref.status)
    np.testing.assert_array_equal(got.positions, ref.positions)


def test_presets_match_adaptive_tracer_on_gpu(hip):
    """test_realtime_reference.py:56-73 with both tracers on the device."""
    from blackhorizon_amd import realtime

    st = orc.Kerr(spin=0.9)
    offsets = np.linspace(0.03, 0.40, 300)
    s0, _ = off.equatorial_fan(st, np.concatenate([offsets, -offsets]))
    hst = hip.KerrSpacetime(spin=0.9)
    truth = hip.trace_rays(hst, s0, escape_radius=60.0, rtol=1e-10)
    for preset in ("low", "medium", "high", "ultra"):
        got = realtime.trace_like_shader(hst, s0, _settings(preset), 60.0, mode="fast")
        assert np.array_equal(_black(got.status), _black(truth.status)), preset


@pytest.mark.parametrize("mode", ["fast", "precise"])
def test_near_extremal_high_preset_on_gpu(hip, mode):
    """test_realtime_reference.py:90-108 with both tracers on the device: at
    a = 0.998 the HIGH preset shades the fan on both sides of the hole
    exactly like the adaptive tracer at rtol 1e-10."""
    from blackhorizon_amd import realtime

    st = orc.Kerr(spin=0.998)
    offsets = np.linspace(0.03, 0.40, 300)
    s0, _ = off.equatorial_fan(st, np.concatenate([offsets, -offsets]))
    hst = hip.KerrSpacetime(spin=0.998)
    truth = hip.trace_rays(hst, s0, escape_radius=60.0, rtol=1e-10)
    got = realtime.trace_like_shader(hst, s0, _settings("high", spin=0.998), 60.0, mode=mode)
    assert np.array_equal(_black(got.status), _black(truth.status))
    # both classes occur on each side (prograde and retrograde edges)
    for side in (slice(0, 300), slice(300, 600)):
        assert 0 < int(_black(got.status[side]).sum()) < 300


def test_schwarzschild_boundary_and_plungers_on_gpu(hip):
    from blackhorizon_amd import realtime

    st = orc.Kerr(spin=0.0)
    s0, impact = off.equatorial_fan(st, np.linspace(0.10, 0.25, 600))
    got = realtime.trace_like_shader(hip.KerrSpacetime(spin=0.0), s0, _settings("medium"), 60.0)
    black = _black(got.status)
    b = 0.5 * (np.abs(impact)[black].max() + np.abs(impact)[~black].min())
    assert abs(b - 3.0 * np.sqrt(3.0)) / (3.0 * np.sqrt(3.0)) < 5e-3
    st = orc.Kerr(spin=0.998)
    s1, _ = off.equatorial_fan(st, np.linspace(0.001, 0.15, 200))
    got = realtime.trace_like_shader(hip.KerrSpacetime(spin=0.998), s1, _settings("low"), 60.0)
    assert np.all(np.isfinite(got.positions)) and np.all(got.status == 1)
