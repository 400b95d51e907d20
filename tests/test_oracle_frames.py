"""Pins the rain-observer frames restatement (oracle/frames_oracle.py) by
closed forms, and the reference's interior-view checks
(tests/test_interior.py:64-103) on the oracle of the shader algorithm; the
GPU tests repeat the latter on the device."""

import math

import numpy as np
import pytest

from oracle import bh_oracle as orc
from oracle import frames_oracle as fr
from oracle import offline_oracle as off

HIGH = off.PRESETS["high"]


@pytest.mark.parametrize("spin", [0.0, 0.9])
@pytest.mark.parametrize("radius", [0.5, 1.9, 2.0, 5.0, 30.0])
def test_rain_observer_and_tetrad_are_normalised(spin, radius):
    st = orc.Kerr(spin=spin)
    pos = np.array([[radius, 0.3, 0.2]])
    u = fr.rain_four_velocity(st, pos)
    assert abs(float(fr.metric_dot(st, pos, u, u)[0]) + 1.0) < 1e-12
    e = fr.build_tetrad(st, pos[0], u[0], np.array([-1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0]))
    gram = np.array([[float(fr.metric_dot(st, pos, e[a][None], e[b][None])[0]) for b in range(4)]
                     for a in range(4)])
    np.testing.assert_allclose(gram, np.diag([-1.0, 1.0, 1.0, 1.0]), atol=1e-11)
    # lower and raise are inverse to each other
    v = np.array([[0.7, -0.2, 0.4, 1.1]])
    np.testing.assert_allclose(fr.raise_index(st, pos, fr.lower_index(st, pos, v)), v,
                               atol=1e-12)


@pytest.mark.parametrize("spin", [0.0, 0.9])
def test_tetrad_rays_are_null_with_unit_camera_frequency(spin):
    st = orc.Kerr(spin=spin)
    s0, mom = fr.rain_camera_states(st, 1.9 if spin == 0.0 else 1.2, 64)
    np.testing.assert_allclose(orc.hamiltonian(st, s0), 0.0, atol=1e-12)
    pos = s0[:, 1:4]
    u = fr.rain_four_velocity(st, pos)
    np.testing.assert_allclose(np.sum(mom * u, axis=-1), 1.0, atol=1e-12)


@pytest.mark.parametrize("radius,expected", [(2.0, 0.5), (8.0, 1.0 / 1.5)])
def test_rain_horizon_overhead_redshift(radius, expected):
    """test_interior.py:106-130: looking radially outward, g = 1 / E =
    1 / (1 + sqrt(2M / r)): exactly 1/2 at the horizon."""
    st = orc.Kerr(spin=0.0)
    pos = np.array([radius, 0.0, 0.0])
    u = fr.rain_four_velocity(st, pos[None, :])[0]
    e = fr.build_tetrad(st, pos, u, np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0]))
    mom = fr.tetrad_ray_momenta(st, pos, e, np.array([[1.0, 0.0, 0.0]]))
    assert abs(1.0 / float(mom[0, 0]) - expected) < 1e-12
    assert abs(expected - 1.0 / (1.0 + math.sqrt(2.0 / radius))) < 1e-15


def _sky_fraction(radius, n=400):
    st = orc.Kerr(spin=0.0)
    s0, _ = fr.rain_camera_states(st, radius, n)
    res = off.trace_like_shader(st, s0, 400.0, *HIGH, interior_stop=0.02)
    return float(np.mean(res.status == 2)), res


def test_interior_sky_checks_on_the_oracle():
    """test_interior.py:64-103 on the shader algorithm's oracle: more than
    half the sky visible at r = 1.9, the sky narrowing toward r = 0.5 but
    above 5% there, and escaping rays with finite, moderate shifts."""
    shallow, _ = _sky_fraction(1.9)
    deep, _ = _sky_fraction(0.5)
    assert shallow > 0.5
    assert 0.05 < deep < shallow
    st = orc.Kerr(spin=0.0)
    s0, mom = fr.rain_camera_states(st, 1.95, 200)
    res = off.trace_like_shader(st, s0, 400.0, *HIGH, interior_stop=0.02)
    esc = res.status == 2
    assert int(esc.sum()) > 100
    assert np.all(mom[esc, 0] > 0.0) and float(mom[esc, 0].max()) < 50.0
