"""DP5(4) step, error ratio and step factor on the device (reference
integrators.py:76-124).

The reference's steppers are generic over a Python ``rhs`` callable.  A HIP
kernel cannot call back into Python, so the device stepper is specialised to
the geodesic right-hand side: pass ``GeodesicRHS(spacetime)`` (or a
``KerrSpacetime``) where the reference passes ``lambda y: geodesic_rhs(st, y)``.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _dev, _lib
from .backend import Array
from .kerr import KerrSpacetime, _mode

# Dormand & Prince (1980) 5(4) tableau (integrators.py:23-54); the kernels
# carry the same constants as literals.
A = (
    (),
    (1.0 / 5.0,),
    (3.0 / 40.0, 9.0 / 40.0),
    (44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0),
    (19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0),
    (9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0),
    (35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0),
)
B5 = A[6] + (0.0,)
E = (71.0 / 57600.0, 0.0, -71.0 / 16695.0, 71.0 / 1920.0, -17253.0 / 339200.0,
     22.0 / 525.0, -1.0 / 40.0)


@dataclass(frozen=True)
class GeodesicRHS:
    """The geodesic right-hand side of one spacetime, as a device stepper key."""

    spacetime: KerrSpacetime
    mode: str = "precise"


def _resolve(rhs) -> GeodesicRHS:
    if isinstance(rhs, GeodesicRHS):
        return rhs
    if isinstance(rhs, KerrSpacetime):
        return GeodesicRHS(rhs)
    raise TypeError("device steppers need GeodesicRHS(spacetime): arbitrary Python "
                    "right-hand sides cannot run inside a HIP kernel")


def _step(rhs, state: Array, h: Array, rtol: float = 1e-9, atol: float = 1e-12):
    g = _resolve(rhs)
    s = _dev.as_device_f64(state, (8,), "state")
    n = s.shape[0]
    hv = _dev.as_device_f64(h, None, "h").reshape(-1)
    if hv.shape[0] != n:
        raise ValueError("h must have shape (n,)")
    new = torch.empty_like(s)
    err = torch.empty_like(s)
    ratio = torch.empty(n, dtype=torch.float64, device=s.device)
    factor = torch.empty(n, dtype=torch.float64, device=s.device)
    lib = _lib.require_device()
    _lib.check(lib.bh_dp5_step_f64(g.spacetime.mass, g.spacetime.spin, _mode(g.mode), rtol, atol,
                                   _lib.ptr(s), _lib.ptr(hv), n, _lib.ptr(new), _lib.ptr(err),
                                   _lib.ptr(ratio), _lib.ptr(factor), _lib.stream_handle()),
               "dormand_prince_step")
    return new, err, ratio, factor


def dormand_prince_step(rhs, state: Array, h: Array):
    """(fifth-order state, embedded error), both (n, 8) (integrators.py:76-107)."""
    new, err, _, _ = _step(rhs, state, h)
    return _dev.like_input(new, state), _dev.like_input(err, state)


def dormand_prince_attempt(rhs, state: Array, h: Array, rtol: float, atol: float):
    """One fused attempt: (new_state, error, ratio, factor) where ratio is
    error_ratio(state, new_state, error, rtol, atol) (integrators.py:110-117)
    and factor is step_factor(ratio) (integrators.py:120-124)."""
    out = _step(rhs, state, h, rtol, atol)
    return tuple(_dev.like_input(t, state) for t in out)
