"""Host prescriptions of the tidal-disruption module (no device): the
reference's own checks (tests/test_dynamics.py:184-209)."""

import numpy as np
import pytest

from blackhorizon_amd import dynamics


def test_tidal_radius_scaling():
    base = dynamics.tidal_radius(1.0, 1e-6, 0.01)
    assert dynamics.tidal_radius(8.0, 1e-6, 0.01) == pytest.approx(2.0 * base)
    assert dynamics.tidal_radius(1.0, 8e-6, 0.01) == pytest.approx(base / 2.0)


def test_hills_mass_marginal_case():
    star_mass, star_radius = 1e-6, 0.02
    m = dynamics.hills_mass(star_mass, star_radius)
    r_t = dynamics.tidal_radius(1.0, star_mass / m, star_radius / m)
    assert abs(r_t - 2.0) < 1e-9


def test_fallback_slope_and_normalization():
    times = np.geomspace(1.0, 1000.0, 4000)
    for partial, slope in ((False, -5.0 / 3.0), (True, -9.0 / 4.0)):
        rate = dynamics.fallback_rate(times, 1.0, 0.5, partial=partial)
        measured = np.log(rate[2000] / rate[1000]) / np.log(times[2000] / times[1000])
        assert abs(measured - slope) < 1e-6
        total = np.trapezoid(rate, times)
        assert abs(total - 0.5) / 0.5 < 0.05
    assert dynamics.fallback_rate(0.5, 1.0, 0.5) == 0.0
