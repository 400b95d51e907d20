"""Synthetic tracing workloads (cameras, spins, tolerances) used by the
benchmark, the parity tests and the fixture generator.

Everything here is host-side parameter bookkeeping; no rays are traced.
The single-camera workloads follow the reference's shadow CLI conventions
(examples/render_shadow.py:56-87): a ``PinholeCamera.from_orbit`` camera,
past-directed null rays, ``escape_radius = 2 * distance`` unless stated.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, asdict

import numpy as np


@dataclass(frozen=True)
class CameraJob:
    """One trace job: a Kerr hole, a pinhole camera and controller settings."""

    spin: float
    distance: float
    inclination_degrees: float
    fov_degrees: float
    width: int
    height: int
    escape_radius: float
    rtol: float
    atol: float
    capture_margin: float = 1e-3
    max_steps: int = 20000
    initial_step: float = 0.1
    max_step: float = 2.0
    min_step: float = 1e-10
    mass: float = 1.0
    azimuth_degrees: float = 0.0

    @property
    def n_rays(self) -> int:
        return self.width * self.height

    @property
    def outer_horizon_radius(self) -> float:
        return self.mass + math.sqrt(self.mass * self.mass - self.spin * self.spin)

    def as_dict(self) -> dict:
        return asdict(self)


def _margin_for_absolute(spin: float, delta: float = 0.01) -> float:
    """Multiplicative capture margin equivalent to r <= r_+ + delta."""
    r_plus = 1.0 + math.sqrt(1.0 - spin * spin)
    return delta / r_plus


def small_frame() -> CameraJob:
    """160x120 frame, a = 0.9, camera at r = 30 M, inclination 60 deg,
    horizontal FOV 30 deg, rtol 1e-8 / atol 1e-10, capture at r_+ + 0.01 M,
    escape at 45 M (the reference's CPU-runnable case)."""
    return CameraJob(spin=0.9, distance=30.0, inclination_degrees=60.0,
                     fov_degrees=30.0, width=160, height=120, escape_radius=45.0,
                     rtol=1e-8, atol=1e-10, capture_margin=_margin_for_absolute(0.9))


def hd_frame(spin: float = 0.9, inclination_degrees: float = 17.0,
             width: int = 1920, height: int = 1080) -> CameraJob:
    """1920x1080 frame at r = 50 M, horizontal FOV 25 deg, rtol 1e-8,
    atol 1e-10, capture at r_+ + 0.01 M, escape at 2 r_obs = 100 M."""
    return CameraJob(spin=spin, distance=50.0, inclination_degrees=inclination_degrees,
                     fov_degrees=25.0, width=width, height=height, escape_radius=100.0,
                     rtol=1e-8, atol=1e-10, capture_margin=_margin_for_absolute(spin))


HD_SPINS = (0.0, 0.5, 0.9, 0.998)
HD_INCLINATIONS = (17.0, 80.0)


def hd_frames() -> list[CameraJob]:
    """The eight 1920x1080 frames: spins x inclinations."""
    return [hd_frame(a, i) for a in HD_SPINS for i in HD_INCLINATIONS]


def shadow_cli_frame() -> CameraJob:
    """Defaults of the reference's shadow CLI (render_shadow.py:29-51,84;
    trace_rays defaults for atol and capture_margin, tracer.py:50-61)."""
    return CameraJob(spin=0.9, distance=60.0, inclination_degrees=85.0,
                     fov_degrees=20.0, width=320, height=240, escape_radius=120.0,
                     rtol=1e-8, atol=1e-12, capture_margin=1e-3)


UNEVEN_MIX_SEED = 355


def uneven_mix(n_jobs: int = 64, seed: int = UNEVEN_MIX_SEED,
               n_min: float = 1e3, n_max: float = 2.6e5) -> list[CameraJob]:
    """Seeded mix of uneven jobs: ray counts log-uniform in [1e3, 2.6e5],
    spin U[-0.998, 0.998], inclination U[5, 85] deg, observer distance
    log-uniform in [20, 1000] M, rtol 10^U[-10, -6], atol = 1e-2 rtol,
    near-square frames, FOV 2 atan(15 / r_obs) so the shadow stays in frame,
    escape at 2 r_obs."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_jobs):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        spin = float(rng.uniform(-0.998, 0.998))
        inc = float(rng.uniform(5.0, 85.0))
        r_obs = float(math.exp(rng.uniform(math.log(20.0), math.log(1000.0))))
        rtol = float(10.0 ** rng.uniform(-10.0, -6.0))
        w = max(1, int(round(math.sqrt(n))))
        h = max(1, int(round(n / w)))
        fov = math.degrees(2.0 * math.atan(15.0 / r_obs))
        jobs.append(CameraJob(spin=spin, distance=r_obs, inclination_degrees=inc,
                              fov_degrees=fov, width=w, height=h,
                              escape_radius=2.0 * r_obs, rtol=rtol, atol=1e-2 * rtol,
                              capture_margin=_margin_for_absolute(spin)))
    return jobs


@dataclass(frozen=True)
class OrbitJob:
    """Timelike orbits of one batch job: positions and coordinate velocities
    (dx^i/dt) of unit-mass particles, integrated over proper time."""

    spin: float
    positions: np.ndarray
    velocities: np.ndarray
    lambda_end: float
    rtol: float
    atol: float = 1e-12
    initial_step: float = 0.05
    max_step: float = 2.0
    min_step: float = 1e-14
    max_steps: int = 10_000_000
    escape_radius: float = 1000.0
    capture_margin: float = 1e-3
    mass: float = 1.0

    @property
    def n_rays(self) -> int:
        return self.positions.shape[0]


ORBIT_SEED = 5


def near_circular_orbits(n: int, spin: float = 0.9, seed: int = ORBIT_SEED, r_min: float = 6.0,
                         r_max: float = 50.0, spread: float = 0.05):
    """Seeded near-circular orbits: radius U[6, 50] M, prograde or retrograde
    with equal probability at the Boyer-Lindquist circular angular velocity
    Omega = +-1 / (r^1.5 +- a) (the reference's test_orbits.py:51-54 and
    118-120), every velocity component scaled by 1 + spread N(0, 1), the
    orbit plane tilted about the x axis by an inclination U[0, 90] deg.
    Positions follow x + i y = (r + i a) e^{i phi} in the plane, velocity
    Omega d/dphi of it.  Returns (positions (n, 3), dx^i/dt (n, 3))."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(r_min, r_max, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    pro = rng.random(n) < 0.5
    a = spin
    omega = np.where(pro, 1.0 / (r ** 1.5 + a), -1.0 / (r ** 1.5 - a))
    x = r * np.cos(phi) - a * np.sin(phi)
    y = r * np.sin(phi) + a * np.cos(phi)
    vx = -omega * y
    vy = omega * x
    v = np.stack([vx, vy, np.zeros(n)], -1) * (1.0 + spread * rng.normal(size=(n, 3)))
    p = np.stack([x, y, np.zeros(n)], -1)
    inc = np.radians(rng.uniform(0.0, 90.0, n))
    c, s = np.cos(inc), np.sin(inc)

    def tilt(u):
        return np.stack([u[:, 0], u[:, 1] * c - u[:, 2] * s, u[:, 1] * s + u[:, 2] * c], -1)

    return tilt(p), tilt(v)


def orbit_batches(n_total: int = 65536, n_jobs: int = 16, spin: float = 0.9,
                  lambda_end: float = 1.0e4, rtol: float = 1e-10,
                  seed: int = ORBIT_SEED) -> list[OrbitJob]:
    """65,536 near-circular timelike orbits around a = 0.9 split into 16
    jobs of seeded uneven sizes (log-uniform weights), each integrated to
    proper time 1e4 M at rtol 1e-10."""
    pos, vel = near_circular_orbits(n_total, spin, seed)
    rng = np.random.default_rng(seed + 1)
    w = np.exp(rng.uniform(0.0, math.log(30.0), n_jobs))
    sizes = np.maximum(1, np.floor(w / w.sum() * n_total)).astype(int)
    sizes[-1] = n_total - sizes[:-1].sum()
    edges = np.concatenate([[0], np.cumsum(sizes)])
    return [OrbitJob(spin=spin, positions=pos[edges[j]:edges[j + 1]],
                     velocities=vel[edges[j]:edges[j + 1]], lambda_end=lambda_end, rtol=rtol)
            for j in range(n_jobs)]
