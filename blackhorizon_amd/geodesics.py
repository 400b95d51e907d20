"""Geodesic dynamics in Kerr-Schild coordinates (reference geodesics.py).

State layout: (n, 8) rows (t, x, y, z, p_t, p_x, p_y, p_z) with covariant
momenta (geodesics.py:3-16).  Every function runs a HIP kernel; NumPy inputs
are copied to the device and results come back as NumPy, tensors stay on the
device.
"""

from __future__ import annotations

import torch

from . import _dev, _lib
from .backend import Array
from .kerr import KerrSpacetime, _mode


def build_state(positions: Array, momenta: Array, t0: float = 0.0) -> Array:
    """(n, 8) state from positions (n, 3) and momenta (n, 4) (geodesics.py:25-39)."""
    if isinstance(positions, torch.Tensor):
        n = positions.shape[0]
        state = torch.empty((n, 8), dtype=positions.dtype, device=positions.device)
        state[:, 0] = t0
        state[:, 1:4] = positions
        state[:, 4:8] = torch.as_tensor(momenta, device=positions.device)
        return state
    import numpy

    positions = numpy.asarray(positions)
    state = numpy.empty((positions.shape[0], 8), dtype=positions.dtype)
    state[:, 0] = t0
    state[:, 1:4] = positions
    state[:, 4:8] = numpy.asarray(momenta)
    return state


def geodesic_rhs(spacetime: KerrSpacetime, state: Array, mode: str = "precise") -> Array:
    """d(state)/d(lambda) from Hamilton's equations (geodesics.py:42-74)."""
    s = _dev.as_device_f64(state, (8,), "state")
    out = torch.empty_like(s)
    lib = _lib.require_device()
    _lib.check(lib.bh_geodesic_rhs_f64(spacetime.mass, spacetime.spin, _mode(mode), _lib.ptr(s),
                                       s.shape[0], _lib.ptr(out), _lib.stream_handle()),
               "geodesic_rhs")
    return _dev.like_input(out, state)


def _diagnostics(spacetime, state, which):
    s = _dev.as_device_f64(state, (8,), "state")
    n = s.shape[0]
    dev = s.device
    bufs = {
        "ham": torch.empty(n, dtype=torch.float64, device=dev) if "ham" in which else None,
        "energy": torch.empty(n, dtype=torch.float64, device=dev) if "energy" in which else None,
        "lz": torch.empty(n, dtype=torch.float64, device=dev) if "lz" in which else None,
        "vel": torch.empty((n, 4), dtype=torch.float64, device=dev) if "vel" in which else None,
    }
    lib = _lib.require_device()
    _lib.check(lib.bh_diagnostics_f64(spacetime.mass, spacetime.spin, _lib.ptr(s), n,
                                      _lib.ptr(bufs["ham"]), _lib.ptr(bufs["energy"]),
                                      _lib.ptr(bufs["lz"]), _lib.ptr(bufs["vel"]),
                                      _lib.stream_handle()), "diagnostics")
    return {k: _dev.like_input(v, state) for k, v in bufs.items() if v is not None}


def hamiltonian(spacetime: KerrSpacetime, state: Array) -> Array:
    """H_ham = 1/2 eta p p - H (l.p)^2, conserved (geodesics.py:77-85)."""
    return _diagnostics(spacetime, state, ("ham",))["ham"]


def conserved_quantities(state: Array, spacetime: KerrSpacetime | None = None):
    """(E, L_z) = (-p_t, x p_y - y p_x) (geodesics.py:88-92)."""
    st = spacetime if spacetime is not None else KerrSpacetime()
    d = _diagnostics(st, state, ("energy", "lz"))
    return d["energy"], d["lz"]


def coordinate_velocity(spacetime: KerrSpacetime, state: Array) -> Array:
    """dx^mu/dlambda = g^{mu nu} p_nu, shape (n, 4) (geodesics.py:95-106)."""
    return _diagnostics(spacetime, state, ("vel",))["vel"]


def null_momentum_from_velocity(spacetime: KerrSpacetime, positions: Array, directions: Array,
                                time_orientation: str = "past", normalize: bool = True) -> Array:
    """Photon covector from a spatial direction (geodesics.py:109-170)."""
    if time_orientation not in ("past", "future"):
        raise ValueError("time_orientation must be 'past' or 'future'")
    pos = _dev.as_device_f64(positions, (3,), "positions")
    dirs = _dev.as_device_f64(directions, (3,), "directions")
    if pos.shape[0] != dirs.shape[0]:
        raise ValueError("positions and directions must have the same length")
    mom = torch.empty((pos.shape[0], 4), dtype=torch.float64, device=pos.device)
    lib = _lib.require_device()
    _lib.check(lib.bh_null_momentum_f64(spacetime.mass, spacetime.spin, _lib.ptr(pos),
                                        _lib.ptr(dirs), pos.shape[0],
                                        1 if time_orientation == "past" else 0,
                                        1 if normalize else 0, _lib.ptr(mom),
                                        _lib.stream_handle()), "null_momentum_from_velocity")
    return _dev.like_input(mom, positions)


def timelike_momentum(spacetime: KerrSpacetime, positions: Array,
                      coordinate_velocities: Array) -> Array:
    """Unit-mass covector from dx^i/dt (geodesics.py:173-209)."""
    pos = _dev.as_device_f64(positions, (3,), "positions")
    vel = _dev.as_device_f64(coordinate_velocities, (3,), "coordinate_velocities")
    if pos.shape[0] != vel.shape[0]:
        raise ValueError("positions and velocities must have the same length")
    n = pos.shape[0]
    mom = torch.empty((n, 4), dtype=torch.float64, device=pos.device)
    bad = torch.zeros(1, dtype=torch.int32, device=pos.device)
    lib = _lib.require_device()
    _lib.check(lib.bh_timelike_momentum_f64(spacetime.mass, spacetime.spin, _lib.ptr(pos),
                                            _lib.ptr(vel), n, _lib.ptr(mom), _lib.ptr(bad),
                                            _lib.stream_handle()), "timelike_momentum")
    if int(bad.item()) > 0:
        raise ValueError("coordinate velocity is not timelike at one or more positions")
    return _dev.like_input(mom, positions)
