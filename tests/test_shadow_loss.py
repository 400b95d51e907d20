"""The shadow-fit loss's pieces (host torch code, no GPU): spheroidal
escape directions, the squared great-circle angle, and the capture margin
(the smooth horizon-proximity term), each pinned by closed forms, by the
Bardeen critical curve and by the oracle's traced capture/escape classes."""

import math

import numpy as np
import pytest
import torch

from blackhorizon_amd import adjoint
from oracle import bh_oracle as orc

f64 = torch.float64


def test_terminal_directions_are_spheroidal_angles():
    """x + i y = (r + i a) sin th e^{i ph}, z = r cos th gives back
    (sin th cos ph, sin th sin ph, cos th) at the Kerr-Schild radius r."""
    rng = np.random.default_rng(1)
    for a in (0.0, 0.7, -0.998):
        r = rng.uniform(1.5, 60.0, 200)
        th = rng.uniform(0.01, math.pi - 0.01, 200)
        ph = rng.uniform(-math.pi, math.pi, 200)
        x = (r * np.cos(ph) - a * np.sin(ph)) * np.sin(th)
        y = (r * np.sin(ph) + a * np.cos(ph)) * np.sin(th)
        z = r * np.cos(th)
        term = np.zeros((200, 8))
        term[:, 1], term[:, 2], term[:, 3] = x, y, z
        n = adjoint.terminal_directions(torch.from_numpy(term), a).numpy()
        want = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
        np.testing.assert_allclose(n, want, atol=1e-13)


def test_angle_sq_is_the_squared_great_circle_angle():
    rng = np.random.default_rng(2)
    u = rng.normal(size=(500, 3))
    v = rng.normal(size=(500, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    psi = np.arctan2(np.linalg.norm(np.cross(u, v), axis=-1), (u * v).sum(-1))
    got = adjoint.angle_sq(torch.from_numpy(u), torch.from_numpy(v)).numpy()
    np.testing.assert_allclose(got, psi ** 2, rtol=1e-12, atol=1e-15)
    # smooth through zero: finite gradient 2 (n - m) at coincident and tiny angles
    for eps in (0.0, 1e-9, 1e-7):
        n = torch.tensor([[1.0, 0.0, 0.0]], dtype=f64, requires_grad=True)
        m = torch.tensor([[math.cos(eps), math.sin(eps), 0.0]], dtype=f64)
        (g,) = torch.autograd.grad(adjoint.angle_sq(n, m).sum(), [n])
        np.testing.assert_allclose(g.numpy(), 2 * (n - m).detach().numpy(), atol=1e-15)


def _camera(spin, width=40, height=1, distance=40.0, inc=60.0, fov=30.0):
    st = orc.Kerr(spin=spin)
    cam = orc.PinholeCamera.from_orbit(distance, inc, fov_degrees=fov, width=width,
                                       height=height)
    return orc.camera_state(st, cam)


def test_capture_margin_schwarzschild_closed_form():
    """a = 0: min_r R / r^4 = 1 - b^2 / 27 with b^2 = |x x p|^2 / E^2."""
    s0 = _camera(0.0)
    m = adjoint.capture_margin(torch.from_numpy(s0), 0.0).numpy()
    b2 = (np.cross(s0[:, 1:4], s0[:, 5:8]) ** 2).sum(-1) / s0[:, 4] ** 2
    np.testing.assert_allclose(m, 1.0 - b2 / 27.0, atol=1e-14)


@pytest.mark.parametrize("a", [0.5, 0.9, -0.9])
def test_radial_potential_vanishes_on_the_critical_curve(a):
    """Bardeen (1973): the constants of spherical photon orbits,
    lam(r) = (r^2 (3 - r) - a^2 (r + 1)) / (a (r - 1)),
    eta(r) = r^3 (4 a^2 - r (r - 3)^2) / (a^2 (r - 1)^2), make the
    potential's minimum zero, at r."""
    aa = abs(a)
    r1 = 2.0 * (1.0 + math.cos(2.0 / 3.0 * math.acos(-aa)))
    r2 = 2.0 * (1.0 + math.cos(2.0 / 3.0 * math.acos(aa)))
    rr = torch.linspace(1.0 + math.sqrt(1 - a * a), 40.0, 400001, dtype=f64)
    for rc in np.linspace(r1 + 1e-3, r2 - 1e-3, 6):
        lam = (rc * rc * (3.0 - rc) - a * a * (rc + 1.0)) / (a * (rc - 1.0))
        eta = rc ** 3 * (4.0 * a * a - rc * (rc - 3.0) ** 2) / (a * a * (rc - 1.0) ** 2)
        v = adjoint._radial_potential(rr, torch.tensor(a, dtype=f64), 1.0,
                                      torch.tensor(lam, dtype=f64), torch.tensor(eta, dtype=f64))
        assert abs(float(v.min())) < 1e-8
        assert abs(float(rr[v.argmin()]) - rc) < 2e-4


@pytest.mark.parametrize("spin", [0.9, -0.5, 0.0])
def test_capture_margin_sign_is_the_traced_class(spin):
    """Positive margin <=> the oracle's trace ends CAPTURED, for every ray of
    a 40 x 30 frame whose margin is not within 1e-3 of the critical curve."""
    s0 = _camera(spin, width=40, height=30, distance=30.0)
    rp = 1.0 + math.sqrt(1.0 - spin * spin)
    res = orc.trace_rays(orc.Kerr(spin=spin), s0, 45.0, rtol=1e-8, atol=1e-10,
                         capture_margin=0.01 / rp)
    m = adjoint.capture_margin(torch.from_numpy(s0), spin).numpy()
    captured = res.status == 1
    assert 100 < captured.sum() < len(m) - 100
    clear = np.abs(m) > 1e-3
    assert np.array_equal((m > 0)[clear], captured[clear])


def test_capture_margin_gradient_matches_central_differences():
    """Envelope theorem at the interior minimiser: autograd of the margin in
    the spin and the initial state equals central differences."""
    s0 = _camera(0.7, width=9, height=1)
    st = torch.from_numpy(s0).requires_grad_(True)
    a = torch.tensor(0.7, dtype=f64, requires_grad=True)
    w = torch.from_numpy(np.random.default_rng(3).normal(size=9))
    gs, ga = torch.autograd.grad((adjoint.capture_margin(st, a) * w).sum(), [st, a])
    f = lambda s, aa: float((adjoint.capture_margin(torch.from_numpy(s), aa) * w).sum())  # noqa
    h = 1e-6
    fd = (f(s0, 0.7 + h) - f(s0, 0.7 - h)) / (2 * h)
    assert abs(fd - float(ga)) <= 1e-6 * abs(fd)
    for i, j in [(0, 5), (3, 6), (8, 7), (4, 1), (2, 3)]:
        d = np.zeros_like(s0)
        d[i, j] = h
        fd = (f(s0 + d, 0.7) - f(s0 - d, 0.7)) / (2 * h)
        assert abs(fd - float(gs[i, j])) <= 1e-6 * abs(fd) + 1e-10, (i, j, fd, float(gs[i, j]))


def test_shadow_loss_terms():
    """Matched rays cost psi^2 of their terminal directions, mismatched rays
    the squared capture margin, other statuses nothing."""
    s0 = torch.from_numpy(_camera(0.7, width=6, height=1))
    term = s0.clone()
    term[:, 1:4] *= 1.1
    status = torch.tensor([2, 1, 2, 1, 3, 2], dtype=torch.int32)
    tstat = torch.tensor([2, 1, 1, 2, 3, 4], dtype=torch.int32)
    dirs = adjoint.terminal_directions(term, 0.7)
    tdirs = torch.roll(dirs, 1, 0)
    loss = adjoint.shadow_loss(term, status, 0.7, tdirs, tstat, state0=s0)
    m = adjoint.capture_margin(s0, 0.7)
    psi2 = adjoint.angle_sq(dirs, tdirs)
    want = psi2[0] + psi2[1] + m[2] ** 2 + m[3] ** 2
    assert abs(float(loss) - float(want)) <= 1e-14 * abs(float(want))
    esc_only = adjoint.shadow_loss(term, status, 0.7, tdirs, tstat, state0=s0, classes=(2,))
    assert abs(float(esc_only) - float(psi2[0])) <= 1e-15


def test_edge_weights():
    """min(1, |m| / scale)^2 of the capture margin, without a graph: at
    a = 0 from the closed form m = 1 - b^2 / 27; rays whose class differs
    from the target's keep weight 1; the weighted loss is the per-ray terms
    scaled by them."""
    s0 = torch.from_numpy(_camera(0.0, width=64, height=1))
    b2 = (np.cross(s0[:, 1:4].numpy(), s0[:, 5:8].numpy()) ** 2).sum(-1) / s0[:, 4].numpy() ** 2
    m = 1.0 - b2 / 27.0
    spin = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    w = adjoint.edge_weights(s0, spin, 0.05)
    assert not w.requires_grad
    np.testing.assert_allclose(w.numpy(), np.minimum(1.0, np.abs(m) / 0.05) ** 2, atol=1e-12)
    assert float(w.min()) < 0.5 and float(w.max()) == 1.0  # the fan crosses the edge
    status = torch.where(torch.from_numpy(m) > 0, 1, 2).to(torch.int32)
    flipped = torch.where(status == 1, 2, 1).to(torch.int32)
    tstat = torch.where(torch.arange(64) % 2 == 0, status, flipped).to(torch.int32)
    w2 = adjoint.edge_weights(s0, 0.0, 0.05, status, tstat)
    mism = (status != tstat).numpy()
    np.testing.assert_array_equal(w2.numpy()[mism], 1.0)
    np.testing.assert_array_equal(w2.numpy()[~mism], w.numpy()[~mism])
    term = s0.clone()
    term[:, 1:4] *= 1.1
    dirs = adjoint.terminal_directions(term, 0.0)
    tdirs = torch.roll(dirs, 3, 0)
    per = [adjoint.shadow_loss(term[i:i + 1], status[i:i + 1], 0.0, tdirs[i:i + 1],
                               tstat[i:i + 1], state0=s0[i:i + 1]) for i in range(64)]
    want = sum(float(w2[i]) * float(per[i]) for i in range(64))
    got = float(adjoint.shadow_loss(term, status, 0.0, tdirs, tstat, state0=s0, ray_weights=w2))
    assert abs(got - want) <= 1e-12 * abs(want)
    for bad in (0.0, -0.1, float("nan")):
        with pytest.raises(ValueError):
            adjoint.edge_weights(s0, 0.0, bad)
