"""Tidal-disruption debris on Kerr geodesics (reference dynamics/tde.py).

The analytic prescriptions (tidal radius, Hills mass, fallback rate, frozen-
in energy spread) are host scalars, as in the reference; the debris
particles' covariant momenta come from the device ``timelike_momentum``
kernel, and their evolution runs in the fused fixed-step kernels
(``integrators.rk4_steps`` for the reference's demo loop,
examples/tde_demo.py:150-190; ``implicit_midpoint_steps`` for long
symplectic runs).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy
import torch

from .backend import Array
from .geodesics import build_state, timelike_momentum
from .kerr import KerrSpacetime


def tidal_radius(bh_mass: float, star_mass: float, star_radius: float) -> float:
    """r_t = R (M / m)^(1/3) (tde.py:62-64)."""
    return star_radius * (bh_mass / star_mass) ** (1.0 / 3.0)


def hills_mass(star_mass: float, star_radius: float) -> float:
    """Largest hole mass that disrupts the star outside r = 2 M (tde.py:67-75)."""
    return (star_radius / 2.0) ** 1.5 / star_mass ** 0.5


def fallback_rate(times, peak_time: float, disrupted_mass: float, partial: bool = False):
    """dM/dt ~ t^(-5/3) (full) or t^(-9/4) (partial) after peak_time,
    normalised to disrupted_mass, zero before (tde.py:78-101)."""
    t = numpy.asarray(times, dtype=float)
    exponent = -9.0 / 4.0 if partial else -5.0 / 3.0
    normalization = disrupted_mass * (-exponent - 1.0) / peak_time
    rate = normalization * (t / peak_time) ** exponent
    return numpy.where(t >= peak_time, rate, 0.0)


def energy_spread_estimate(bh_mass: float, star_radius: float, r_t: float) -> float:
    """Frozen-in specific energy spread M R / r_t^2 (tde.py:188-192)."""
    return bh_mass * star_radius / r_t ** 2


def marginally_bound_speed(spacetime: KerrSpacetime, position, direction) -> float:
    """Coordinate speed along ``direction`` giving E = -p_t = 1 at
    ``position``: 80 bisections on [1e-6, 0.999999] (tde.py:27-59); a trial
    speed outside the light cone counts as unbound."""
    position = numpy.asarray(position, dtype=float)
    direction = numpy.asarray(direction, dtype=float)

    def energy(speed: float) -> float:
        try:
            p = timelike_momentum(spacetime, position[None, :], (speed * direction)[None, :])
        except ValueError:
            return numpy.inf
        return float(-p[0, 0])

    low, high = 1e-6, 0.999999
    if energy(low) >= 1.0:
        raise ValueError("even a static particle is unbound here")
    for _ in range(80):
        mid = 0.5 * (low + high)
        if energy(mid) < 1.0:
            low = mid
        else:
            high = mid
    return 0.5 * (low + high)


@dataclass(frozen=True)
class DebrisStream:
    """Initial conditions of a tidally disrupted star (tde.py:104-118)."""

    states: Array
    specific_energies: Array
    bound_fraction: float
    tidal_radius: float


def generate_debris_stream(spacetime: KerrSpacetime, star_mass: float, star_radius: float,
                           n_particles: int = 2000, penetration: float = 1.0,
                           prograde: bool = True, seed: int = 0,
                           device_states: bool = False) -> DebrisStream:
    """A star disrupted at pericenter r_p = r_t / penetration into test
    particles (tde.py:121-185): positions sample the stellar sphere (seeded
    normal offsets of scale R/2, those beyond R pulled inside), all share
    the centre's marginally bound tangential velocity, momenta from
    timelike_momentum.  ``device_states`` keeps the states as a device
    tensor (for the fused steppers); otherwise NumPy, as the reference."""
    r_t = tidal_radius(1.0, star_mass, star_radius)
    r_p = r_t / penetration
    if r_p < 4.0 * spacetime.outer_horizon_radius:
        raise ValueError("pericenter too close to the horizon for the test-particle "
                         "prescription; reduce penetration or star compactness")
    rng = numpy.random.default_rng(seed)
    offsets = rng.normal(scale=star_radius / 2.0, size=(n_particles, 3))
    radii = numpy.linalg.norm(offsets, axis=-1)
    keep = radii <= star_radius
    offsets[~keep] *= (star_radius / numpy.maximum(radii[~keep], 1e-12) * 0.99)[:, None]
    center = numpy.array([r_p, 0.0, 0.0])
    direction = numpy.array([0.0, 1.0 if prograde else -1.0, 0.0])
    speed = marginally_bound_speed(spacetime, center, direction)
    positions = center[None, :] + offsets
    velocities = numpy.tile(speed * direction, (n_particles, 1))
    pos = torch.from_numpy(positions)
    vel = torch.from_numpy(velocities)
    if device_states:
        from . import _dev

        pos = pos.to(_dev.device())
        vel = vel.to(_dev.device())
        momenta = timelike_momentum(spacetime, pos, vel)
        states = build_state(pos, momenta)
        energies = -states[:, 4]
        bound = float((energies < 1.0).double().mean().item())
    else:
        momenta = timelike_momentum(spacetime, positions, velocities)
        states = build_state(positions, momenta)
        energies = -states[:, 4]
        bound = float(numpy.mean(energies < 1.0))
    return DebrisStream(states=states, specific_energies=energies, bound_fraction=bound,
                        tidal_radius=r_t)
