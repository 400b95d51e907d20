"""Kerr spacetime in Cartesian Kerr-Schild coordinates (reference kerr.py).

``KerrSpacetime`` keeps the reference's constructor checks and analytic radii
(kerr.py:61-112, host scalars).  The pointwise geometry -- KS radius r, the
scalar H = M r^3 / (r^4 + a^2 z^2), the null covector l and their analytic
gradients (kerr.py:118-193) -- is evaluated by the HIP kernel behind
``bh_geometry_f64``; the tracer kernels inline the same device code.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy
import torch

from . import _dev, _lib
from .backend import Array


@dataclass(frozen=True)
class KerrSchildGeometry:
    """kerr.py:31-49: radius, h (...), l (..., 3), grad_h (..., 3),
    grad_l (..., 3, 3) with grad_l[..., i, j] = d l_j / d x_i."""

    radius: Array
    h: Array
    l: Array
    grad_h: Array | None = None
    grad_l: Array | None = None


class KerrSpacetime:
    """A Kerr black hole of mass M and spin a = J / M (kerr.py:52-112)."""

    def __init__(self, mass: float = 1.0, spin: float = 0.0) -> None:
        if mass <= 0.0:
            raise ValueError("mass must be positive")
        if abs(spin) > mass:
            raise ValueError("Kerr bound violated: |spin| must be <= mass")
        self.mass = float(mass)
        self.spin = float(spin)

    @property
    def outer_horizon_radius(self) -> float:
        m, a = self.mass, self.spin
        return m + math.sqrt(m * m - a * a)

    @property
    def inner_horizon_radius(self) -> float:
        m, a = self.mass, self.spin
        return m - math.sqrt(m * m - a * a)

    def ergosphere_radius(self, theta: float) -> float:
        m, a = self.mass, self.spin
        return m + math.sqrt(m * m - a * a * math.cos(theta) ** 2)

    def isco_radius(self, prograde: bool = True) -> float:
        """Bardeen, Press & Teukolsky (1972) ISCO radius."""
        chi = abs(self.spin) / self.mass
        z1 = 1.0 + (1.0 - chi * chi) ** (1.0 / 3.0) * (
            (1.0 + chi) ** (1.0 / 3.0) + (1.0 - chi) ** (1.0 / 3.0))
        z2 = math.sqrt(3.0 * chi * chi + z1 * z1)
        root = math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))
        return self.mass * (3.0 + z2 + (-1.0 if prograde else 1.0) * root)

    def photon_orbit_radius(self, prograde: bool = True) -> float:
        """Equatorial circular photon orbit (Bardeen 1973)."""
        chi = abs(self.spin) / self.mass
        sign = -1.0 if prograde else 1.0
        return 2.0 * self.mass * (1.0 + math.cos((2.0 / 3.0) * math.acos(sign * chi)))

    # -- pointwise geometry on the device ---------------------------------

    def geometry(self, x: Array, y: Array, z: Array, gradients: bool = False,
                 mode: str = "precise") -> KerrSchildGeometry:
        """Kerr-Schild geometry at points (x, y, z), broadcastable arrays
        (kerr.py:118-193): radius and h on the broadcast shape, l and grad_h
        on shape + (3,), grad_l on shape + (3, 3).  The points are flattened
        for the device and the results reshaped back; NumPy in -> NumPy out,
        tensor in -> tensor out."""
        ref = x if isinstance(x, (numpy.ndarray, torch.Tensor)) else (
            y if isinstance(y, (numpy.ndarray, torch.Tensor)) else z)
        dev = _dev.device()
        comps = [v.detach().to(device=dev, dtype=torch.float64) if isinstance(v, torch.Tensor)
                 else torch.as_tensor(numpy.asarray(v, dtype=numpy.float64)).to(dev)
                 for v in (x, y, z)]
        shape = torch.broadcast_shapes(*(c.shape for c in comps))
        xyz = torch.stack([c.expand(shape).reshape(-1) for c in comps], dim=-1).contiguous()
        n = xyz.shape[0]
        lib = _lib.require_device()
        radius = torch.empty(n, dtype=torch.float64, device=dev)
        h = torch.empty(n, dtype=torch.float64, device=dev)
        l = torch.empty((n, 3), dtype=torch.float64, device=dev)
        gh = torch.empty((n, 3), dtype=torch.float64, device=dev) if gradients else None
        gl = torch.empty((n, 3, 3), dtype=torch.float64, device=dev) if gradients else None
        _lib.check(lib.bh_geometry_f64(self.mass, self.spin, _mode(mode), _lib.ptr(xyz), n,
                                       _lib.ptr(radius), _lib.ptr(h), _lib.ptr(l), _lib.ptr(gh),
                                       _lib.ptr(gl), _lib.stream_handle()), "geometry")
        tails = ((), (), (3,), (3,), (3, 3))
        out = [None if t is None else _dev.like_input(t.reshape(tuple(shape) + tail), ref)
               for t, tail in zip((radius, h, l, gh, gl), tails)]
        return KerrSchildGeometry(radius=out[0], h=out[1], l=out[2], grad_h=out[3], grad_l=out[4])

    def kerr_schild_radius(self, x: Array, y: Array, z: Array) -> Array:
        """kerr.py:114-116 (broadcastable arrays, as geometry)."""
        return self.geometry(x, y, z).radius

    def metric(self, positions: Array) -> Array:
        """g_{mu nu} = eta + 2 H l_mu l_nu at positions of shape (..., 4)
        (kerr.py:201-217); result (..., 4, 4)."""
        return self._metric(positions, inverse=False)

    def inverse_metric(self, positions: Array) -> Array:
        """g^{mu nu} = eta - 2 H l^mu l^nu at (..., 4) (kerr.py:219-235)."""
        return self._metric(positions, inverse=True)

    def _metric(self, positions: Array, inverse: bool) -> Array:
        pos = positions.detach().to(device=_dev.device(), dtype=torch.float64) \
            if isinstance(positions, torch.Tensor) else \
            torch.as_tensor(numpy.asarray(positions, dtype=numpy.float64)).to(_dev.device())
        if pos.dim() < 1 or pos.shape[-1] != 4:
            raise ValueError(f"positions must have shape (..., 4), got {tuple(pos.shape)}")
        lead = tuple(pos.shape[:-1])
        flat = pos.reshape(-1, 4)
        geo = self.geometry(flat[:, 1], flat[:, 2], flat[:, 3])
        l4 = torch.cat([(-1.0 if inverse else 1.0) * torch.ones_like(geo.h)[:, None], geo.l], -1)
        eta = torch.diag(torch.tensor([-1.0, 1.0, 1.0, 1.0], dtype=torch.float64,
                                      device=pos.device)).expand(flat.shape[0], 4, 4)
        sign = -2.0 if inverse else 2.0
        g = eta + sign * geo.h[:, None, None] * (l4[:, :, None] * l4[:, None, :])
        return _dev.like_input(g.reshape(lead + (4, 4)), positions)


def _mode(mode: str) -> int:
    if mode == "precise":
        return _lib.MODE_PRECISE
    if mode == "fast":
        return _lib.MODE_FAST
    raise ValueError("mode must be 'fast' or 'precise'")
