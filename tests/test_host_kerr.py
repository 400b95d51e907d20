"""The package's host-side KerrSpacetime scalars (kerr.py:61-112): the
reference's analytic-radius checks (tests/test_kerr.py:20-49) on this class,
and, at intermediate spins where no limit pins them, the defining equations
each radius solves (horizons r^2 - 2Mr + a^2 = 0; static limit
r^2 - 2Mr + a^2 cos^2(theta) = 0; ISCO r^2 - 6Mr -+ 8a sqrt(Mr) - 3a^2 = 0,
Bardeen, Press & Teukolsky 1972; circular photon orbit
r^(3/2) - 3M r^(1/2) -+ 2a sqrt(M) = 0), so a wrong sign or a swapped
prograde branch fails.  Host code only: no device is used."""

import math

import pytest

from blackhorizon_amd.kerr import KerrSpacetime


def test_horizons():
    assert KerrSpacetime(spin=0.0).outer_horizon_radius == pytest.approx(2.0)
    assert KerrSpacetime(spin=1.0).outer_horizon_radius == pytest.approx(1.0)
    st = KerrSpacetime(spin=0.6)
    assert st.outer_horizon_radius == pytest.approx(1.8)
    assert st.inner_horizon_radius == pytest.approx(0.2)


def test_isco_limits():
    assert KerrSpacetime(spin=0.0).isco_radius() == pytest.approx(6.0)
    assert KerrSpacetime(spin=1.0).isco_radius(prograde=True) == pytest.approx(1.0)
    assert KerrSpacetime(spin=1.0).isco_radius(prograde=False) == pytest.approx(9.0)


def test_photon_orbit_limits():
    assert KerrSpacetime(spin=0.0).photon_orbit_radius() == pytest.approx(3.0)
    assert KerrSpacetime(spin=1.0).photon_orbit_radius(True) == pytest.approx(1.0)
    assert KerrSpacetime(spin=1.0).photon_orbit_radius(False) == pytest.approx(4.0)


def test_ergosphere():
    st = KerrSpacetime(spin=0.9)
    assert st.ergosphere_radius(0.0) == pytest.approx(st.outer_horizon_radius)
    assert st.ergosphere_radius(math.pi / 2.0) == pytest.approx(2.0)


def test_kerr_bound_enforced():
    with pytest.raises(ValueError):
        KerrSpacetime(spin=1.01)
    with pytest.raises(ValueError):
        KerrSpacetime(mass=-1.0)
    with pytest.raises(ValueError):
        KerrSpacetime(mass=0.0)


@pytest.mark.parametrize("mass", [1.0, 2.5])
@pytest.mark.parametrize("chi", [0.0, 0.3, 0.5, 0.9, 0.998])
def test_radii_solve_their_defining_equations(mass, chi):
    a = chi * mass
    st = KerrSpacetime(mass=mass, spin=a)
    for r in (st.outer_horizon_radius, st.inner_horizon_radius):
        assert abs(r * r - 2.0 * mass * r + a * a) < 1e-12 * mass * mass
    assert st.inner_horizon_radius <= st.outer_horizon_radius
    for theta in (0.0, 0.4, 1.1, math.pi / 2.0):
        r = st.ergosphere_radius(theta)
        assert abs(r * r - 2.0 * mass * r + (a * math.cos(theta)) ** 2) < 1e-12 * mass * mass
        assert r >= st.outer_horizon_radius - 1e-12
    for prograde, sign in ((True, 1.0), (False, -1.0)):
        r = st.isco_radius(prograde)
        residual = r * r - 6.0 * mass * r + sign * 8.0 * a * math.sqrt(mass * r) - 3.0 * a * a
        assert abs(residual) < 1e-10 * mass * mass
        r = st.photon_orbit_radius(prograde)
        residual = r ** 1.5 - 3.0 * mass * math.sqrt(r) + sign * 2.0 * a * math.sqrt(mass)
        assert abs(residual) < 1e-10 * mass ** 1.5
    # prograde orbits sit inside retrograde ones, outside the horizon
    assert st.outer_horizon_radius <= st.photon_orbit_radius(True) <= st.isco_radius(True)
    assert st.photon_orbit_radius(True) <= st.photon_orbit_radius(False)
    assert st.isco_radius(True) <= st.isco_radius(False)
    if chi == 0.0:
        assert st.isco_radius(True) == pytest.approx(6.0 * mass)
        assert st.isco_radius(False) == pytest.approx(6.0 * mass)


def test_negative_spin_mirrors_the_prograde_branch():
    """The radii depend on |a| (kerr.py:92-112): the same numbers for a and -a."""
    for a in (0.3, 0.9):
        p, m = KerrSpacetime(spin=a), KerrSpacetime(spin=-a)
        assert p.isco_radius(True) == m.isco_radius(True)
        assert p.photon_orbit_radius(False) == m.photon_orbit_radius(False)
        assert p.outer_horizon_radius == m.outer_horizon_radius
