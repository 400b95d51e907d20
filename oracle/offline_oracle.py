"""CPU oracle of the reference's offline tile tracer (test infrastructure).

Restates, in NumPy and in the reference's expression order, the tracing
part of the offline renderer: the per-tile adaptive DP5(4) loop with
classification before each step and equatorial disk crossings
(offline/render.py:184-309), the crossing refinement by bisection along
RK4 sub-steps (render.py:312-349), the flat-space escape direction
(render.py:164-181), the photon-shell retrace of saturated rays
(render.py:612-630), the camera sample grid (render.py:101-162,
realtime/fly_camera.py:27-107) and the real-time shader's fixed-step
algorithm (realtime/reference.py:41-188).  Only tests, smoke() and the
CPU-baseline leg of bench.py use it, as the checker; the product never
imports it.

Pinning (tests/test_oracle_offline.py): Schwarzschild face-on disk hits
against an independent quadrature of the orbit equation; crossings at
z = 0 inside the annulus; the trace loop without a disk against the
tracer restatement (bh_oracle.trace_rays) where the two loops must agree;
flat directions of rays escaping radially; saturated/retrace bookkeeping
on constructed cases; for the shader algorithm the reference's own accuracy
checks (every preset classifies an a = 0.9 fan as the adaptive tracer does,
b_c = 3 sqrt(3) within 0.5%, no numerical failures) and disk hits on the
crossing chord.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from oracle.bh_oracle import (Kerr, RayStatus, dormand_prince_step, error_ratio, geodesic_rhs,
                              rk4_step, step_factor)

_WORLD_UP = np.array([0.0, 0.0, 1.0])
_PITCH_LIMIT = math.radians(89.0)


@dataclass
class FlyCamera:
    """Position + yaw/pitch camera (realtime/fly_camera.py:20-107)."""

    position: np.ndarray
    yaw: float = math.pi
    pitch: float = 0.0

    @classmethod
    def from_orbit(cls, distance, inclination_degrees, azimuth_degrees=0.0):
        inc = math.radians(inclination_degrees)
        az = math.radians(azimuth_degrees)
        position = np.array([distance * math.sin(inc) * math.cos(az),
                             distance * math.sin(inc) * math.sin(az),
                             distance * math.cos(inc)])
        cam = cls(position=position)
        cam.look_at(np.zeros(3))
        return cam

    @property
    def distance_from_origin(self) -> float:
        return float(np.linalg.norm(self.position))

    def basis(self):
        cp = math.cos(self.pitch)
        forward = np.array([cp * math.cos(self.yaw), cp * math.sin(self.yaw),
                            math.sin(self.pitch)])
        right = np.cross(forward, _WORLD_UP)
        norm = np.linalg.norm(right)
        if norm < 1e-9:
            right = np.array([math.sin(self.yaw), -math.cos(self.yaw), 0.0])
        else:
            right = right / norm
        up = np.cross(right, forward)
        return forward, right, up

    def look_at(self, target):
        offset = np.asarray(target, dtype=float) - self.position
        norm = np.linalg.norm(offset)
        if norm < 1e-12:
            return
        d = offset / norm
        self.yaw = math.atan2(d[1], d[0])
        self.pitch = max(-_PITCH_LIMIT, min(_PITCH_LIMIT, math.asin(float(d[2]))))


def pinhole_grid(width, height, fov_degrees, supersample):
    """render.py:101-144 (tan(fov / 2) scales the horizontal coordinate;
    order: pixel row, column, subpixel row, column)."""
    tan_half = math.tan(math.radians(fov_degrees) / 2.0)
    aspect = width / height
    ss = supersample
    offsets = (np.arange(ss) + 0.5) / ss
    px = (np.arange(width)[:, None] + offsets[None, :]).reshape(-1)
    py = (np.arange(height)[:, None] + offsets[None, :]).reshape(-1)
    u = px / width * 2.0 - 1.0
    v = 1.0 - py / height * 2.0
    u_grid = np.tile(u.reshape(width, ss)[None, :, None, :], (height, 1, ss, 1))
    v_grid = np.tile(v.reshape(height, ss)[:, None, :, None], (1, width, 1, ss))
    local = np.stack([np.ones(u_grid.size), u_grid.reshape(-1) * tan_half,
                      v_grid.reshape(-1) * (tan_half / aspect)], axis=-1)
    return local / np.linalg.norm(local, axis=-1, keepdims=True)


def subpixel_directions(camera, width, height, fov_degrees, supersample):
    """render.py:147-161."""
    forward, right, up = camera.basis()
    local = pinhole_grid(width, height, fov_degrees, supersample)
    return (local[:, 0:1] * forward[None, :] + local[:, 1:2] * right[None, :]
            + local[:, 2:3] * up[None, :])


def flat_directions(st: Kerr, positions, momenta, p_t=None):
    """render.py:164-181."""
    geo = st.geometry(positions[:, 0], positions[:, 1], positions[:, 2])
    energies = 1.0 if p_t is None else p_t
    lp = -energies + np.sum(geo.l * momenta, axis=-1)
    directions = momenta - 2.0 * (geo.h * lp)[:, None] * geo.l
    norm = np.sqrt(np.sum(directions * directions, axis=-1))
    return directions / norm[:, None]


@dataclass
class TileOutput:
    status: np.ndarray
    escape_directions: np.ndarray
    hit_positions: np.ndarray
    hit_momenta: np.ndarray
    saturated: np.ndarray
    steps: np.ndarray


def refine_crossings(st: Kerr, rhs, y0, h_full, disk_radii):
    """render.py:312-349: 40 bisections of the crossing fraction along one
    RK4 sub-step from the pre-step state."""
    z0 = y0[:, 3]
    low = np.zeros(y0.shape[0])
    high = np.ones(y0.shape[0])
    for _ in range(40):
        mid = 0.5 * (low + high)
        y_mid = rk4_step(rhs, y0, h_full * mid)
        same_side = y_mid[:, 3] * z0 > 0.0
        low = np.where(same_side, mid, low)
        high = np.where(same_side, high, mid)
    y_cross = rk4_step(rhs, y0, h_full * 0.5 * (low + high))
    r_hit = st.kerr_schild_radius(y_cross[:, 1], y_cross[:, 2], y_cross[:, 3])
    hits = (r_hit >= disk_radii[0]) & (r_hit <= disk_radii[1])
    return hits, y_cross[hits][:, 1:4], y_cross[hits][:, 5:8]


def trace_tile(st: Kerr, state0, escape_radius, disk_radii, rtol, max_steps, max_step=2.0,
               capture_margin=1e-3, interior_stop=None, factor_fn=None):
    """render.py:184-309 (settings reduced to max_step and capture_margin)."""
    sf = factor_fn or step_factor
    n = state0.shape[0]
    if interior_stop is None:
        capture_radius = st.outer_horizon_radius * (1.0 + capture_margin)
        capture_status = int(RayStatus.CAPTURED)
    else:
        capture_radius = interior_stop
        capture_status = int(RayStatus.TERMINATED)

    def rhs(batch):
        return geodesic_rhs(st, batch)

    states = state0.copy()
    status = np.full((n,), int(RayStatus.IN_FLIGHT), dtype=np.int32)
    steps = np.zeros((n,), dtype=np.int64)
    h = np.full((n,), 0.1)
    hit_positions = np.zeros((n, 3))
    hit_momenta = np.zeros((n, 3))
    iterations = 0
    while iterations < 4 * max_steps:
        idx = np.nonzero(status == int(RayStatus.IN_FLIGHT))[0]
        if idx.size == 0:
            break
        iterations += 1
        y = states[idx]
        radius = st.kerr_schild_radius(y[:, 1], y[:, 2], y[:, 3])
        captured = radius <= capture_radius
        if interior_stop is not None:
            momentum_norm = np.sqrt(np.sum(y[:, 5:8] * y[:, 5:8], axis=-1))
            captured = captured | (momentum_norm > 1.0e3)
            captured = captured | ((radius <= capture_radius * 1.002) & (momentum_norm > 50.0))
        escaped = ~captured & (radius >= escape_radius)
        exhausted = steps[idx] >= max_steps
        status[idx[captured]] = capture_status
        status[idx[escaped]] = int(RayStatus.ESCAPED)
        status[idx[exhausted & ~captured & ~escaped]] = int(RayStatus.MAX_STEPS)
        alive = ~(captured | escaped | exhausted)
        idx = idx[alive]
        if idx.size == 0:
            continue
        y = states[idx]
        h_a = h[idx]
        with np.errstate(all="ignore"):
            y_new, err = dormand_prince_step(rhs, y, h_a)
            ratio = error_ratio(y, y_new, err, rtol, rtol * 1e-3)
        accept = ratio <= 1.0
        if disk_radii is not None:
            crossed = accept & (y[:, 3] * y_new[:, 3] < 0.0)
            if bool(crossed.any()):
                with np.errstate(all="ignore"):
                    hits, positions, momenta = refine_crossings(st, rhs, y[crossed], h_a[crossed],
                                                                disk_radii)
                targets = idx[np.nonzero(crossed)[0][hits]]
                status[targets] = int(RayStatus.DISK)
                hit_positions[targets] = positions
                hit_momenta[targets] = momenta
        still = status[idx] == int(RayStatus.IN_FLIGHT)
        y = np.where(accept[:, None], y_new, y)
        states[idx] = np.where(still[:, None], y, states[idx])
        steps[idx] = steps[idx] + (accept & still).astype(steps.dtype)
        h[idx] = np.clip(h_a * sf(ratio), 0.0, max_step)
    saturated = (status == int(RayStatus.MAX_STEPS)) | (steps >= int(0.85 * max_steps))
    escape_mask = status == int(RayStatus.ESCAPED)
    directions = np.zeros((n, 3))
    if bool(escape_mask.any()):
        directions[escape_mask] = flat_directions(st, states[escape_mask][:, 1:4],
                                                  states[escape_mask][:, 5:8],
                                                  p_t=states[escape_mask][:, 4])
    return TileOutput(status=status, escape_directions=directions, hit_positions=hit_positions,
                      hit_momenta=hit_momenta, saturated=saturated, steps=steps)


def trace_tile_retrace(st: Kerr, state0, escape_radius, disk_radii, rtol, refine_rtol,
                       max_steps, max_step=2.0, capture_margin=1e-3, interior_stop=None,
                       factor_fn=None):
    """The first pass plus the photon-shell retrace (render.py:592-630):
    saturated rays that neither hit the disk nor terminated are traced
    again at refine_rtol with four times the step budget, and their status,
    escape direction and disk hit are replaced (the saturated flag and the
    step counts stay the first pass's)."""
    kw = dict(max_step=max_step, capture_margin=capture_margin, interior_stop=interior_stop,
              factor_fn=factor_fn)
    result = trace_tile(st, state0, escape_radius, disk_radii, rtol, max_steps, **kw)
    retrace = result.saturated & (result.status != int(RayStatus.DISK)) & (
        result.status != int(RayStatus.TERMINATED))
    if bool(retrace.any()):
        refined = trace_tile(st, state0[np.nonzero(retrace)[0]], escape_radius, disk_radii,
                             refine_rtol, max_steps * 4, **kw)
        result.status[retrace] = refined.status
        result.escape_directions[retrace] = refined.escape_directions
        result.hit_positions[retrace] = refined.hit_positions
        result.hit_momenta[retrace] = refined.hit_momenta
    return result, retrace


# ---------------------------------------------------------------------------
# the real-time shader's algorithm (realtime/reference.py:41-188)
# ---------------------------------------------------------------------------

PRESETS = {  # realtime/settings.py:33-50 (max_steps, step_scale)
    "low": (256, 0.35), "medium": (512, 0.22), "high": (1024, 0.14), "ultra": (2048, 0.08)}


@dataclass
class ShaderOutput:
    status: np.ndarray
    positions: np.ndarray
    momenta: np.ndarray
    hit_radii: np.ndarray | None
    hit_positions: np.ndarray | None
    hit_momenta: np.ndarray | None


def trace_like_shader(st: Kerr, state0, escape_radius, max_steps=512, step_scale=0.22,
                      min_step=1e-4, max_step=4.0, capture_margin=2e-2, momentum_bailout=1e3,
                      disk_radii=None, interior_stop=None):
    """realtime/reference.py:41-188 with the RenderSettings fields as arguments."""
    r_plus = st.outer_horizon_radius
    if interior_stop is None:
        capture_radius = r_plus * (1.0 + capture_margin)
        capture_status = int(RayStatus.CAPTURED)
    else:
        capture_radius = interior_stop
        capture_status = int(RayStatus.TERMINATED)
    n = state0.shape[0]
    state = state0.copy()
    status = np.full((n,), int(RayStatus.MAX_STEPS), dtype=np.int32)
    active = np.ones((n,), dtype=bool)
    hit_radii = np.full((n,), np.nan)
    hit_positions = np.full((n, 3), np.nan)
    hit_momenta = np.full((n, 4), np.nan)
    if interior_stop is not None:
        max_steps = min(max_steps, 1024)
    for _ in range(max_steps):
        idx = np.nonzero(active)[0]
        if idx.size == 0:
            break
        y = state[idx]
        radius = st.kerr_schild_radius(y[:, 1], y[:, 2], y[:, 3])
        momentum2 = np.sum(y[:, 5:8] * y[:, 5:8], axis=-1)
        captured = (radius <= capture_radius) | (momentum2 >= momentum_bailout ** 2)
        if interior_stop is not None:
            captured = captured | ((radius <= capture_radius * 1.002) & (momentum2 > 2500.0))
        escaped = ~captured & (radius >= escape_radius)
        status[idx[captured]] = capture_status
        status[idx[escaped]] = int(RayStatus.ESCAPED)
        active[idx] = ~(captured | escaped)
        keep = ~(captured | escaped)
        idx = idx[keep]
        if idx.size == 0:
            break
        y = y[keep]
        radius = radius[keep]
        if interior_stop is None:
            scale_length = radius - r_plus
        else:
            scale_length = np.minimum(np.maximum(np.abs(radius - r_plus), 0.15 * radius),
                                      0.5 * radius)
        h = np.clip(step_scale * scale_length, min_step, max_step)
        h = np.minimum(h, 1.0 / np.maximum(1.0, np.sqrt(momentum2[keep])))
        hh = h[:, None]
        k1 = geodesic_rhs(st, y)
        k2 = geodesic_rhs(st, y + 0.5 * hh * k1)
        k3 = geodesic_rhs(st, y + 0.5 * hh * k2)
        k4 = geodesic_rhs(st, y + hh * k3)
        y_new = y + (hh / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        if disk_radii is not None:
            crossed = y[:, 3] * y_new[:, 3] < 0.0
            if bool(crossed.any()):
                t_cross = y[crossed, 3] / (y[crossed, 3] - y_new[crossed, 3])
                interpolated = y[crossed] + t_cross[:, None] * (y_new[crossed] - y[crossed])
                r_hit = st.kerr_schild_radius(interpolated[:, 1], interpolated[:, 2],
                                              interpolated[:, 3])
                on_disk = (r_hit >= disk_radii[0]) & (r_hit <= disk_radii[1])
                hit_local = np.zeros(y.shape[0], dtype=bool)
                hit_local[np.nonzero(crossed)[0][on_disk]] = True
                targets = idx[hit_local]
                status[targets] = int(RayStatus.DISK)
                active[targets] = False
                hit_radii[targets] = r_hit[on_disk]
                hit_positions[targets] = interpolated[on_disk][:, 1:4]
                hit_momenta[targets] = interpolated[on_disk][:, 4:8]
        state[idx] = y_new
    disk = disk_radii is not None
    return ShaderOutput(status=status, positions=state[:, 1:4].copy(),
                        momenta=state[:, 4:8].copy(), hit_radii=hit_radii if disk else None,
                        hit_positions=hit_positions if disk else None,
                        hit_momenta=hit_momenta if disk else None)


def equatorial_fan(st: Kerr, offsets, distance=30.0):
    """test_realtime_reference.py:25-38: past-directed equatorial rays from
    (D, 0, 0) along (-1, offset, 0); returns (states, impact parameters)."""
    n = offsets.shape[0]
    positions = np.tile([[distance, 0.0, 0.0]], (n, 1))
    directions = np.stack([-np.ones(n), offsets, np.zeros(n)], axis=-1)
    directions /= np.linalg.norm(directions, axis=-1, keepdims=True)
    from oracle.bh_oracle import build_state, conserved_quantities, null_momentum_from_velocity
    state0 = build_state(positions, null_momentum_from_velocity(st, positions, directions, "past"))
    energy, l_z = conserved_quantities(state0)
    return state0, l_z / energy
