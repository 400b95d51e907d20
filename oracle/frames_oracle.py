"""Observer frames for interior cameras: a NumPy restatement of the
reference's rain observer and tetrad camera (frames.py).

TEST INFRASTRUCTURE ONLY (see oracle/bh_oracle.py for the import rule): it
builds the initial states of the reference's interior-view checks
(tests/test_interior.py:41-130), which the tests then trace on the device.

Pinned by closed forms in tests/test_oracle_frames.py: g(u, u) = -1 for the
rain observer, g(e_a, e_b) = eta_ab for the tetrad, null traced momenta of
unit camera frequency, and the rain observer's factor-2 redshift of
overhead starlight at r = 2M, g = 1 / (1 + sqrt(2M / r))
(test_interior.py:106-130).
"""

from __future__ import annotations

import numpy as np

from .bh_oracle import Kerr


def _geo(st: Kerr, positions):
    return st.geometry(positions[:, 0], positions[:, 1], positions[:, 2])


def metric_dot(st: Kerr, positions, a_vec, b_vec):
    """g(A, B) of contravariant 4-vectors, g = eta + 2 H l l with covariant
    l = (1, l_i) (frames.py:33-60)."""
    geo = _geo(st, positions)
    eta = -a_vec[:, 0] * b_vec[:, 0] + np.sum(a_vec[:, 1:4] * b_vec[:, 1:4], axis=-1)
    l_a = a_vec[:, 0] + np.sum(geo.l * a_vec[:, 1:4], axis=-1)
    l_b = b_vec[:, 0] + np.sum(geo.l * b_vec[:, 1:4], axis=-1)
    return eta + 2.0 * geo.h * l_a * l_b


def lower_index(st: Kerr, positions, vectors):
    """v_mu = g_mu_nu v^nu (frames.py:63-77)."""
    geo = _geo(st, positions)
    l_v = vectors[:, 0] + np.sum(geo.l * vectors[:, 1:4], axis=-1)
    out = np.empty_like(vectors)
    out[:, 0] = -vectors[:, 0] + 2.0 * geo.h * l_v
    out[:, 1:4] = vectors[:, 1:4] + (2.0 * geo.h * l_v)[:, None] * geo.l
    return out


def raise_index(st: Kerr, positions, covectors):
    """v^mu = g^mu_nu v_nu with g^-1 = eta - 2 H l^ l^, l^ = (-1, l_i)
    (frames.py:80-99)."""
    geo = _geo(st, positions)
    l_up_v = -covectors[:, 0] + np.sum(geo.l * covectors[:, 1:4], axis=-1)
    out = np.empty_like(covectors)
    out[:, 0] = -covectors[:, 0] + 2.0 * geo.h * l_up_v
    out[:, 1:4] = covectors[:, 1:4] - (2.0 * geo.h * l_up_v)[:, None] * geo.l
    return out


def rain_four_velocity(st: Kerr, positions):
    """u^mu of the rain observer, free fall from rest at infinity: covariant
    u = (-1, w l) with w = -sqrt(2H) / (1 + sqrt(2H)), raised
    (frames.py:102-126)."""
    geo = _geo(st, positions)
    root = np.sqrt(np.maximum(2.0 * geo.h, 0.0))
    w = -root / (1.0 + root)
    cov = np.empty(positions.shape[:-1] + (4,))
    cov[:, 0] = -1.0
    cov[:, 1:4] = w[:, None] * geo.l
    return raise_index(st, positions, cov)


def build_tetrad(st: Kerr, position, four_velocity, forward, up):
    """Orthonormal tetrad (e0 = u, e1 = view, e2 = forward x up, e3 = up) by
    Gram-Schmidt in the metric (frames.py:129-181)."""
    pos = np.asarray(position, dtype=float)[None, :]

    def dot(a, b):
        return float(metric_dot(st, pos, a[None, :], b[None, :])[0])

    e0 = np.asarray(four_velocity, dtype=float)
    norm0 = dot(e0, e0)
    if norm0 >= 0.0:
        raise ValueError("observer 4-velocity must be timelike")
    e0 = e0 / np.sqrt(-norm0)
    seeds = [np.concatenate([[0.0], np.asarray(forward, dtype=float)]),
             np.concatenate([[0.0], np.cross(forward, up).astype(float)]),
             np.concatenate([[0.0], np.asarray(up, dtype=float)])]
    basis = [e0]
    for seed in seeds:
        v = seed + dot(seed, e0) * e0
        for spatial in basis[1:]:
            v = v - dot(seed, spatial) * spatial
        norm = dot(v, v)
        if norm <= 1e-14:
            raise ValueError("degenerate tetrad seed directions")
        basis.append(v / np.sqrt(norm))
    return np.stack(basis)


def tetrad_ray_momenta(st: Kerr, position, tetrad, local_directions):
    """Past-directed traced covariant momenta of camera rays looking along
    local directions n: the physical photon e0 - n^i e_i (unit camera
    frequency), negated and lowered (frames.py:184-218)."""
    n = local_directions.shape[0]
    spatial = (local_directions[:, 0:1] * tetrad[1][None, :]
               + local_directions[:, 1:2] * tetrad[2][None, :]
               + local_directions[:, 2:3] * tetrad[3][None, :])
    physical = tetrad[0][None, :] - spatial
    positions = np.tile(np.asarray(position, dtype=float)[None, :], (n, 1))
    return lower_index(st, positions, -physical)


def fibonacci_sphere(n: int):
    """Nearly uniform unit directions (n, 3) (test_interior.py:27-38)."""
    i = np.arange(n, dtype=float) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / n)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * i
    return np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)],
                    axis=-1)


def rain_camera_states(st: Kerr, radius: float, n_rays: int, forward=(-1.0, 0.0, 0.0),
                       up=(0.0, 0.0, 1.0)):
    """(n_rays, 8) states of a rain observer at (radius, 0, 0) seeing the
    whole sky (Fibonacci directions), as test_interior.py:41-58 builds them;
    also the traced momenta."""
    pos = np.array([radius, 0.0, 0.0])
    u = rain_four_velocity(st, pos[None, :])[0]
    tetrad = build_tetrad(st, pos, u, np.array(forward), np.array(up))
    mom = tetrad_ray_momenta(st, pos, tetrad, fibonacci_sphere(n_rays))
    state = np.concatenate([np.zeros((n_rays, 1)), np.tile(pos[None, :], (n_rays, 1)), mom], -1)
    return state, mom
