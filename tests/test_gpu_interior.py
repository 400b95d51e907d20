"""GPU: the interior-camera modes against the reference's own interior
checks.  Cameras are rain observers (free fall from rest at infinity,
frames.py:102-218, restated in oracle/frames_oracle.py) inside or at the
Schwarzschild horizon, their whole sky traced backwards:

* the shader algorithm (realtime/reference.py:41-188 with interior_stop):
  more than half the sky shows the outside universe just inside the horizon
  (r = 1.9), the sky narrows deeper in but stays above 5% at r = 0.5, and
  escaping rays carry finite, moderate energies -- test_interior.py:64-103
  -- with every status equal to the oracle's bit for bit;
* the offline tile tracer (render.py:184-309 with interior_stop): 64 x 40
  frames from r = 1.5 see the outside universe looking outward, a mix of sky
  and stop radius looking sideways, the stop radius looking inward; they are
  deterministic and match the offline oracle -- test_offline.py:311-345,
  whose emission stage is out of scope.
"""

import numpy as np
import pytest

from oracle import bh_oracle as orc
from oracle import frames_oracle as fr
from oracle import offline_oracle as off

pytestmark = pytest.mark.gpu


def _high():
    from blackhorizon_amd.realtime import QualityPreset, RenderSettings

    return RenderSettings().apply_preset(QualityPreset.HIGH)


def _device_sky(hip, radius, n):
    from blackhorizon_amd import realtime

    st = orc.Kerr(spin=0.0)
    s0, mom = fr.rain_camera_states(st, radius, n)
    got = realtime.trace_like_shader(hip.KerrSpacetime(spin=0.0), s0, _high(), 400.0,
                                     interior_stop=0.02)
    ref = off.trace_like_shader(st, s0, 400.0, *off.PRESETS["high"], interior_stop=0.02)
    np.testing.assert_array_equal(got.status, ref.status)
    np.testing.assert_array_equal(got.positions, ref.positions)
    return got, mom


def test_interior_sky_visible_and_narrowing(hip):
    shallow, _ = _device_sky(hip, 1.9, 400)
    deep, _ = _device_sky(hip, 0.5, 400)
    f_shallow = float(np.mean(shallow.status == 2))
    f_deep = float(np.mean(deep.status == 2))
    print(f"sky fraction: r = 1.9 {f_shallow:.3f}, r = 0.5 {f_deep:.3f}")
    assert f_shallow > 0.5
    assert 0.05 < f_deep < f_shallow
    # the others end at the stop radius or by diverging momentum, or run out
    # of the 1024 interior passes (reference.py:101-103)
    assert set(np.unique(shallow.status).tolist()) <= {2, 3, 6}


def test_interior_escaping_rays_have_finite_shift(hip):
    got, mom = _device_sky(hip, 1.95, 200)
    esc = got.status == 2
    assert int(esc.sum()) > 100
    assert np.all(mom[esc, 0] > 0.0) and float(mom[esc, 0].max()) < 50.0
    assert np.all(np.isfinite(got.momenta[esc]))


@pytest.mark.parametrize("view", ["outward", "sideways", "inward"])
def test_offline_interior_frame_sees_the_sky(hip, view):
    """From r = 1.5: looking outward (FOV 80 deg) every pixel sees the sky;
    looking sideways with a 120 deg FOV the frame's edge toward the
    singularity ends at the stop radius (TERMINATED) and the rest escapes;
    looking inward (80 deg) every ray ends at the stop radius.  Statuses and
    step counts equal the offline oracle's, two launches agree bit for
    bit."""
    forward, fov = {"outward": ((1.0, 0.0, 0.0), 80.0), "sideways": ((0.0, 1.0, 0.0), 120.0),
                    "inward": ((-1.0, 0.0, 0.0), 80.0)}[view]
    from blackhorizon_amd import offline

    st = orc.Kerr(spin=0.0)
    radius, width, height = 1.5, 64, 40
    pos = np.array([radius, 0.0, 0.0])
    u = fr.rain_four_velocity(st, pos[None, :])[0]
    tetrad = fr.build_tetrad(st, pos, u, np.array(forward), np.array([0.0, 0.0, 1.0]))
    half_w = np.tan(np.radians(fov) / 2.0)
    half_h = half_w * height / width
    cols = ((np.arange(width) + 0.5) / width * 2.0 - 1.0) * half_w
    rows = (1.0 - (np.arange(height) + 0.5) / height * 2.0) * half_h
    v, uu = np.meshgrid(rows, cols, indexing="ij")
    local = np.stack([np.ones_like(uu), uu, v], -1).reshape(-1, 3)
    local /= np.linalg.norm(local, axis=-1, keepdims=True)
    mom = fr.tetrad_ray_momenta(st, pos, tetrad, local)
    n = local.shape[0]
    s0 = np.concatenate([np.zeros((n, 1)), np.tile(pos, (n, 1)), mom], -1)
    assert st.kerr_schild_radius(pos[None, 0], pos[None, 1], pos[None, 2])[0] < 2.0
    kw = dict(escape_radius=100.0, disk_radii=None, rtol=1e-9, max_steps=3000, interior_stop=0.05)
    ref = off.trace_tile(st, s0, **kw)
    args = (hip.KerrSpacetime(spin=0.0), s0, 100.0, offline.OfflineSettings(), None, 1e-9, 3000)
    first = offline.trace_tile(*args, interior_stop=0.05, mode="precise")
    second = offline.trace_tile(*args, interior_stop=0.05, mode="precise")
    for name in ("status", "steps", "escape_directions"):
        np.testing.assert_array_equal(getattr(first, name), getattr(second, name))
    np.testing.assert_array_equal(first.status, ref.status)
    np.testing.assert_array_equal(first.steps, ref.steps)
    # flat escape directions (render.py:164-181), from final states that
    # carry the device's last-bit rounding (measured up to 8e-14)
    np.testing.assert_allclose(first.escape_directions, ref.escape_directions, rtol=0, atol=1e-11)
    esc = first.status == 2
    counts = np.bincount(first.status, minlength=7)
    print(f"offline interior frame ({view}): statuses {counts.tolist()}")
    assert np.all(np.isfinite(first.escape_directions))
    assert set(np.unique(first.status).tolist()) <= {2, 3, 6}
    if view == "outward":
        assert esc.all()
    elif view == "sideways":
        assert counts[2] > 0.05 * n and counts[6] > 0.05 * n
    else:
        assert counts[6] == n
