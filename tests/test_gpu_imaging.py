"""GPU: camera rays built on the device and the Stage-1 image shading, against
the oracle and against the reference-held shadow_gpu.png."""

import numpy as np
import pytest
import torch

from conftest import golden
from oracle import bh_oracle as orc
from blackhorizon_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("frame", ["small", "shadow_cli"])
def test_device_camera_states_bitwise(hip, frame):
    job = workloads.small_frame() if frame == "small" else workloads.shadow_cli_frame()
    kw = dict(fov_degrees=job.fov_degrees, width=job.width, height=job.height)
    cam = hip.PinholeCamera.from_orbit(job.distance, job.inclination_degrees, **kw)
    got = cam.device_states(hip.KerrSpacetime(spin=job.spin)).cpu().numpy()
    ocam = orc.PinholeCamera.from_orbit(job.distance, job.inclination_degrees, **kw)
    np.testing.assert_array_equal(got, orc.camera_state(orc.Kerr(spin=job.spin), ocam))


def test_render_image_matches_oracle_shading(hip):
    """Same traced states shaded on the device and by the oracle; every third
    ray is taken from a trace capped at 30 steps so that MAX_STEPS pixels
    appear beside captured and escaped ones."""
    from blackhorizon_amd import imaging
    from blackhorizon_amd.tracer import TraceResult

    job = workloads.small_frame()
    st = hip.KerrSpacetime(spin=job.spin)
    cam = hip.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees, width=job.width,
                                       height=job.height)
    s0 = cam.device_states(st)
    kw = dict(rtol=job.rtol, atol=job.atol, capture_margin=job.capture_margin)
    full = hip.trace_rays(st, s0, job.escape_radius, max_steps=job.max_steps, **kw)
    capped = hip.trace_rays(st, s0, job.escape_radius, max_steps=30, **kw)
    pick = (torch.arange(s0.shape[0], device=s0.device) % 3 == 0)
    res = TraceResult(states=torch.where(pick[:, None], capped.states, full.states),
                      status=torch.where(pick, capped.status, full.status),
                      steps=full.steps, iterations=0)
    status = res.status.cpu().numpy()
    assert {1, 2, 3} <= set(np.unique(status))
    img = imaging.render_image(st, res, job.width, job.height)
    ref = orc.render_image(orc.Kerr(spin=job.spin), res.states.cpu().numpy(), status,
                           job.width, job.height)
    np.testing.assert_array_equal(img, ref)


def test_shadow_gpu_png_rendered_on_device(hip):
    """The render_shadow defaults (a = 0.9, inclination 85 deg, D = 60,
    FOV 20 deg, 320x240, rtol 1e-8) rendered entirely on the device: the
    reference-held image's shadow pixels exactly, its checker cells on
    >= 99.9% of escaped pixels."""
    from blackhorizon_amd import imaging

    cls = golden("shadow_gpu_classes.npz")["classes"]
    job = workloads.shadow_cli_frame()
    st = hip.KerrSpacetime(spin=job.spin)
    cam = hip.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees, width=job.width,
                                       height=job.height)
    res = hip.trace_rays(st, cam.device_states(st), job.escape_radius, max_steps=job.max_steps,
                         rtol=job.rtol, ray_order=cam.tile_order(8))
    img = imaging.render_image(st, res, job.width, job.height)
    got = np.full(cls.shape, 3)
    got[np.all(img == (0, 0, 0), axis=-1)] = 0
    got[np.all(img == (24, 30, 52), axis=-1)] = 1
    got[np.all(img == (104, 128, 178), axis=-1)] = 2
    np.testing.assert_array_equal(got == 0, cls == 0)
    esc = cls != 0
    assert np.all(got[esc] != 3)
    assert np.mean(got[esc] == cls[esc]) > 0.999


def test_celestial_checkerboard_matches_oracle(hip):
    """celestial_checkerboard (imaging.py:27-50) on the device equals the
    oracle on 20000 random unit directions, the cell centres and the poles
    and axes (whose cell indices fall on boundaries), for several cell
    counts."""
    from blackhorizon_amd import imaging
    from test_oracle_imaging import cell_centres

    rng = np.random.default_rng(11)
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    axes = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [-1, -0.0, 0], [0, 1, 0],
                     [0, -1, 0], [0, 0, 7.5], [0, 0, -3.0]], dtype=np.float64)
    for cells in (1, 5, 12, 31):
        dirs = np.concatenate([d, cell_centres(cells)[0], axes])
        got = imaging.celestial_checkerboard(dirs, cells)
        assert got.dtype == np.uint8 and got.shape == dirs.shape
        np.testing.assert_array_equal(got, orc.celestial_checkerboard(dirs, cells))
    t = torch.from_numpy(d).cuda()
    np.testing.assert_array_equal(imaging.celestial_checkerboard(t),
                                  orc.celestial_checkerboard(d))
    assert imaging.celestial_checkerboard(np.zeros((0, 3))).shape == (0, 3)
