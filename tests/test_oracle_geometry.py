"""Pins the oracle's Kerr-Schild geometry and RHS with closed forms,
invariants and finite differences (re-derived from the reference's
test_kerr.py and test_gradients.py; none of these retype the formulas)."""

import math

import numpy as np
import pytest

from oracle import bh_oracle as orc


def random_points(n, r_min=2.5, r_max=40.0, seed=3):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    return rng.uniform(r_min, r_max, size=(n, 1)) * u


def test_horizons_isco_photon_orbit():
    assert orc.Kerr(spin=0.0).outer_horizon_radius == pytest.approx(2.0)
    assert orc.Kerr(spin=1.0).outer_horizon_radius == pytest.approx(1.0)
    st = orc.Kerr(spin=0.6)
    assert st.outer_horizon_radius == pytest.approx(1.8)
    assert st.inner_horizon_radius == pytest.approx(0.2)
    assert orc.Kerr(spin=0.0).isco_radius() == pytest.approx(6.0)
    assert orc.Kerr(spin=1.0).isco_radius(True) == pytest.approx(1.0)
    assert orc.Kerr(spin=1.0).isco_radius(False) == pytest.approx(9.0)
    assert orc.Kerr(spin=0.0).photon_orbit_radius() == pytest.approx(3.0)
    assert orc.Kerr(spin=1.0).photon_orbit_radius(True) == pytest.approx(1.0)
    assert orc.Kerr(spin=1.0).photon_orbit_radius(False) == pytest.approx(4.0)


def test_kerr_bound():
    with pytest.raises(ValueError):
        orc.Kerr(spin=1.01)
    with pytest.raises(ValueError):
        orc.Kerr(mass=-1.0)


def test_radius_closed_forms():
    st = orc.Kerr(spin=0.8)
    z = np.array([1.5, 3.0, -7.0])
    np.testing.assert_allclose(st.kerr_schild_radius(np.zeros(3), np.zeros(3), z), np.abs(z),
                               rtol=1e-14)
    x = np.array([2.0, 5.0, 11.0])
    np.testing.assert_allclose(st.kerr_schild_radius(x, np.zeros(3), np.zeros(3)),
                               np.sqrt(x * x - 0.64), rtol=1e-14)
    pts = random_points(50)
    r = orc.Kerr(spin=0.0).kerr_schild_radius(pts[:, 0], pts[:, 1], pts[:, 2])
    np.testing.assert_allclose(r, np.linalg.norm(pts, axis=-1), rtol=1e-13)


def test_radius_solves_quartic():
    st = orc.Kerr(spin=0.9)
    pts = random_points(200, 1.0, 50.0, seed=5)
    x, y, z = pts.T
    r = st.kerr_schild_radius(x, y, z)
    rho2 = x * x + y * y + z * z
    resid = r ** 4 - (rho2 - 0.81) * r ** 2 - 0.81 * z * z
    np.testing.assert_allclose(resid / r ** 4, 0.0, atol=1e-13)


def test_null_covector_and_metric_inverse():
    st = orc.Kerr(spin=0.95)
    pts = random_points(200)
    geo = st.geometry(pts[:, 0], pts[:, 1], pts[:, 2])
    np.testing.assert_allclose(-1.0 + np.sum(geo.l * geo.l, axis=-1), 0.0, atol=1e-12)
    pos = np.concatenate([np.zeros((200, 1)), pts], axis=-1)
    ident = np.einsum("nab,nbc->nac", st.metric(pos), st.inverse_metric(pos))
    np.testing.assert_allclose(ident, np.broadcast_to(np.eye(4), ident.shape), atol=1e-12)


def test_schwarzschild_limit():
    st = orc.Kerr(spin=0.0)
    pts = random_points(100)
    r = np.linalg.norm(pts, axis=-1)
    geo = st.geometry(pts[:, 0], pts[:, 1], pts[:, 2], gradients=True)
    np.testing.assert_allclose(geo.h, 1.0 / r, rtol=1e-13)
    np.testing.assert_allclose(geo.l, pts / r[:, None], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(geo.grad_h, -pts / (r ** 3)[:, None], rtol=1e-12)
    # d(x_j / r)/dx_i = (delta_ij - l_i l_j) / r
    l = pts / r[:, None]
    expect = (np.eye(3)[None] - l[:, :, None] * l[:, None, :]) / r[:, None, None]
    np.testing.assert_allclose(geo.grad_l, expect, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("spin", [0.9, -0.5, 0.998])
def test_gradients_match_central_differences(spin):
    st = orc.Kerr(spin=spin)
    pts = random_points(300, r_min=2.2, r_max=60.0, seed=11)
    geo = st.geometry(pts[:, 0], pts[:, 1], pts[:, 2], gradients=True)
    eps = 1e-6
    for axis in range(3):
        d = np.zeros(3)
        d[axis] = eps
        gp = st.geometry(*(pts + d).T)
        gm = st.geometry(*(pts - d).T)
        np.testing.assert_allclose(geo.grad_h[:, axis], (gp.h - gm.h) / (2 * eps),
                                   rtol=1e-5, atol=1e-10)
        np.testing.assert_allclose(geo.grad_l[:, axis, :], (gp.l - gm.l) / (2 * eps),
                                   rtol=1e-5, atol=1e-9)


def _random_states(st, n, seed, timelike=False):
    rng = np.random.default_rng(seed)
    pos = random_points(n, 2.5, 30.0, seed=seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    mom = orc.null_momentum_from_velocity(st, pos, d, "past")
    return orc.build_state(pos, mom)


@pytest.mark.parametrize("spin", [0.0, 0.7, -0.998])
def test_rhs_is_hamiltons_equations(spin):
    """RHS = (dH/dp, -dH/dx) by central differences of the Hamiltonian."""
    st = orc.Kerr(spin=spin)
    s = _random_states(st, 64, 7)
    f = orc.geodesic_rhs(st, s)
    eps = 1e-6
    for j in range(8):
        d = np.zeros(8)
        d[j] = eps
        dh = (orc.hamiltonian(st, s + d) - orc.hamiltonian(st, s - d)) / (2 * eps)
        if j < 4:  # position -> -dp/dlambda (t: stationary, dp_t = 0)
            np.testing.assert_allclose(f[:, 4 + j], -dh, rtol=1e-6, atol=1e-8)
        else:      # momentum -> dx/dlambda
            np.testing.assert_allclose(f[:, j - 4], dh, rtol=1e-6, atol=1e-8)
    assert np.all(f[:, 4] == 0.0)


def test_rhs_conserves_hamiltonian():
    st = orc.Kerr(spin=0.9)
    s = _random_states(st, 64, 9)
    f = orc.geodesic_rhs(st, s)
    eps = 1e-6
    dh = (orc.hamiltonian(st, s + eps * f) - orc.hamiltonian(st, s - eps * f)) / (2 * eps)
    np.testing.assert_allclose(dh, 0.0, atol=1e-8)


def test_null_momentum_is_null_and_tangential_b():
    st = orc.Kerr(spin=0.9)
    s = _random_states(st, 100, 3)
    np.testing.assert_allclose(orc.hamiltonian(st, s), 0.0, atol=1e-12)
    assert np.allclose(np.abs(s[:, 4]), 1.0)
    assert np.all(s[:, 4] > 0)  # past-directed and normalised: p_t = +1
    sch = orc.Kerr(spin=0.0)
    mom = orc.null_momentum_from_velocity(sch, np.array([[3.0, 0.0, 0.0]]),
                                          np.array([[0.0, 1.0, 0.0]]), "future")
    e, lz = orc.conserved_quantities(orc.build_state(np.array([[3.0, 0.0, 0.0]]), mom))
    np.testing.assert_allclose(lz / e, 3.0 * math.sqrt(3.0), rtol=1e-12)


def test_timelike_momentum_normalised():
    st = orc.Kerr(spin=0.9)
    r = 6.0
    om = 1.0 / (r ** 1.5 + 0.9)
    pos = np.array([[r, 0.9, 0.0]])
    mom = orc.timelike_momentum(st, pos, np.array([[-om * 0.9, om * r, 0.0]]))
    np.testing.assert_allclose(orc.hamiltonian(st, orc.build_state(pos, mom)), -0.5, rtol=1e-13)
    with pytest.raises(ValueError):
        orc.timelike_momentum(st, pos, np.array([[0.0, 1.5, 0.0]]))


def test_camera_basis_and_fov():
    cam = orc.PinholeCamera.from_orbit(30.0, 60.0, fov_degrees=30.0, width=160, height=120)
    for v in (cam.forward, cam.right, cam.up):
        assert abs(np.linalg.norm(v) - 1.0) < 1e-15
    assert abs(cam.forward @ cam.right) < 1e-15 and abs(cam.forward @ cam.up) < 1e-15
    np.testing.assert_allclose(cam.forward, -np.asarray(cam.position) / 30.0, atol=1e-15)
    d = cam.ray_directions().reshape(120, 160, 3)
    # leftmost/rightmost pixel centres sit at +-(1 - 1/W) of the horizontal half-FOV
    half = math.tan(math.radians(15.0)) * (1.0 - 1.0 / 160)
    ang = math.atan(half)
    mid_row = 0.5 * (d[59] + d[60])
    for col, sign in ((0, -1), (159, 1)):
        v = mid_row[col] / np.linalg.norm(mid_row[col])
        assert abs(math.atan2(v @ cam.right, v @ cam.forward) - sign * ang) < 1e-12
    assert (d[0, 80] @ cam.up) > 0 and (d[119, 80] @ cam.up) < 0   # row 0 is the top
