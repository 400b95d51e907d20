"""Pins the autograd derivative oracle: its forward agrees with the NumPy
oracle (statuses, accepted steps) and its gradients with central finite
differences of its own terminal points."""

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


@pytest.mark.parametrize("which", ["state", "spin"])
def test_gradients_match_central_differences(which):
    s0 = _rays(0.7, n=4, seed=2)
    w = torch.from_numpy(np.random.default_rng(3).normal(size=(4, 8)))
    w[:, 0] = 0.0  # coordinate time is not part of a direction loss

    def f(state, spin):
        y, st, _ = ado.trace_terminal(state, spin, **KW)
        return (y * w).sum(), st

    st0 = torch.from_numpy(s0).requires_grad_(True)
    sp0 = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    L, status = f(st0, sp0)
    gs, ga = torch.autograd.grad(L, [st0, sp0])
    eps = 1e-6
    with torch.no_grad():
        if which == "spin":
            lp, _ = f(torch.from_numpy(s0), torch.tensor(0.7 + eps, dtype=torch.float64))
            lm, _ = f(torch.from_numpy(s0), torch.tensor(0.7 - eps, dtype=torch.float64))
            fd = (lp - lm).item() / (2 * eps)
            assert abs(fd - ga.item()) <= 1e-5 * abs(fd) + 1e-9
        else:
            for i, j in [(0, 5), (1, 6), (2, 7), (3, 1), (0, 3)]:
                d = np.zeros_like(s0)
                d[i, j] = eps
                lp, _ = f(torch.from_numpy(s0 + d), torch.tensor(0.7, dtype=torch.float64))
                lm, _ = f(torch.from_numpy(s0 - d), torch.tensor(0.7, dtype=torch.float64))
                fd = (lp - lm).item() / (2 * eps)
                assert abs(fd - gs[i, j].item()) <= 1e-5 * abs(fd) + 1e-7, (i, j, fd, gs[i, j])
