"""Pins the autograd derivative oracle: its forward agrees with the NumPy
oracle (statuses, accepted steps); its gradients equal central differences
of the same discrete map (accepted step sizes held fixed) and, on escaped
rays, central differences of the full adaptive trace."""

import numpy as np
import pytest
import torch

from oracle import ad_oracle as ado
from oracle import bh_oracle as orc


def _rays(spin, n=6, seed=0):
    rng = np.random.default_rng(seed)
    cam = orc.PinholeCamera.from_orbit(40.0, 45.0, fov_degrees=30.0, width=n, height=1)
    st = orc.Kerr(spin=spin)
    s0 = orc.camera_state(st, cam)
    s0[:, 6] += rng.normal(scale=0.01, size=n)  # vary the aim out of the row
    return s0


KW = dict(escape_radius=45.0, rtol=1e-8, atol=1e-10, capture_margin=0.01 / 1.714)


def test_forward_matches_numpy_oracle():
    s0 = _rays(0.7, n=16, seed=1)
    ref = orc.trace_rays(orc.Kerr(spin=0.7), s0, **KW)
    y, status, steps = ado.trace_terminal(torch.from_numpy(s0), torch.tensor(0.7,
                                          dtype=torch.float64), **KW)
    np.testing.assert_array_equal(status.numpy(), ref.status)
    np.testing.assert_array_equal(steps.numpy(), ref.steps)
    esc = ref.status == 2
    r = np.linalg.norm(y.detach().numpy()[esc, 1:4], axis=-1)
    assert np.all(np.abs(r - 45.0) < 0.5)  # terminal points sit on the escape sphere


def test_gradients_equal_fixed_step_central_differences():
    s0 = _rays(0.7, n=4, seed=2)
    w = torch.from_numpy(np.random.default_rng(3).normal(size=(4, 8)))
    w[:, 0] = 0.0
    st0 = torch.from_numpy(s0).requires_grad_(True)
    sp0 = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    y, status, _, hl = ado.trace_terminal(st0, sp0, return_steps=True, **KW)
    assert set(status.tolist()) == {1, 2}
    gs, ga = torch.autograd.grad((y * w).sum(), [st0, sp0])
    kw = dict(escape_radius=KW["escape_radius"], capture_margin=KW["capture_margin"])
    with torch.no_grad():
        base = ado.terminal_fixed_steps(torch.from_numpy(s0), 0.7, hl, **kw)
        np.testing.assert_allclose(base.numpy(), y.detach().numpy(), rtol=1e-12, atol=1e-12)
        f = lambda s, a: float((ado.terminal_fixed_steps(s, a, hl, **kw) * w).sum())  # noqa
        eps = 1e-6
        for i, j in [(0, 5), (1, 6), (2, 7), (3, 1), (0, 3), (2, 2)]:
            d = np.zeros_like(s0)
            d[i, j] = eps
            fd = (f(torch.from_numpy(s0 + d), 0.7) - f(torch.from_numpy(s0 - d), 0.7)) / (2 * eps)
            assert abs(fd - gs[i, j].item()) <= 1e-6 * abs(fd) + 1e-8, (i, j, fd, gs[i, j])
        sa = lambda a: f(torch.from_numpy(s0), torch.tensor(a, dtype=torch.float64))  # noqa
        fd = (sa(0.7 + eps) - sa(0.7 - eps)) / (2 * eps)
        assert abs(fd - ga.item()) <= 1e-6 * abs(fd)


def test_escaped_ray_gradients_match_adaptive_central_differences():
    """On escaped rays the controller's own response is O(rtol)."""
    s0 = _rays(0.7, n=4, seed=2)
    w = torch.from_numpy(np.random.default_rng(3).normal(size=(4, 8)))
    w[:, 0] = 0.0
    st0 = torch.from_numpy(s0).requires_grad_(True)
    y, status, _ = ado.trace_terminal(st0, torch.tensor(0.7, dtype=torch.float64), **KW)
    (gs,) = torch.autograd.grad((y * w).sum(), [st0])
    esc = np.nonzero(status.numpy() == 2)[0]
    eps = 1e-6
    with torch.no_grad():
        f = lambda s: float((ado.trace_terminal(s, torch.tensor(0.7, dtype=torch.float64),  # noqa
                                                **KW)[0] * w).sum())
        for i in esc:
            for j in (3, 5, 6):
                d = np.zeros_like(s0)
                d[i, j] = eps
                fd = (f(torch.from_numpy(s0 + d)) - f(torch.from_numpy(s0 - d))) / (2 * eps)
                assert abs(fd - gs[i, j].item()) <= 1e-5 * abs(fd) + 1e-8, (i, j, fd, gs[i, j])
