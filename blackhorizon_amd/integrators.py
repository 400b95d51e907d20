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


def error_ratio(state: Array, new_state: Array, error: Array, rtol: float, atol: float, *,
                mode: str = "precise") -> Array:
    """Scaled RMS error over the last axis (integrators.py:110-117): a value
    <= 1 accepts the step; non-finite values become +inf.  The three arrays
    share one shape (..., d) with 1 <= d <= 128; the result has shape (...).
    The precise build matches NumPy's rounding and summation order bit for
    bit."""
    s = _dev.as_device_f64(state)
    sn = _dev.as_device_f64(new_state)
    se = _dev.as_device_f64(error)
    if s.dim() == 0 or sn.shape != s.shape or se.shape != s.shape:
        raise ValueError("state, new_state and error must share one shape (..., d)")
    d = s.shape[-1]
    if not 1 <= d <= 128:
        raise ValueError("the last axis must hold 1 to 128 values")
    lead = s.shape[:-1]
    n = s.numel() // d
    ratio = torch.empty(n, dtype=torch.float64, device=s.device)
    lib = _lib.require_device()
    _lib.check(lib.bh_error_ratio_f64(_mode(mode), _lib.ptr(s), _lib.ptr(sn), _lib.ptr(se), n, d,
                                      float(rtol), float(atol), _lib.ptr(ratio),
                                      _lib.stream_handle()), "error_ratio")
    return _dev.like_input(ratio.reshape(lead), state)


def step_factor(ratio: Array, *, mode: str = "precise") -> Array:
    """Step-size multiplier clip(0.9 max(ratio, 1e-16)^-0.2, 0.2, 5)
    (integrators.py:120-124); the precise build rounds the power correctly."""
    is_t = isinstance(ratio, torch.Tensor)
    r = ratio.detach().to(_dev.device(), torch.float64).contiguous() if is_t else \
        _dev.as_device_f64(ratio)
    shape = r.shape
    r = r.reshape(-1)
    out = torch.empty_like(r)
    lib = _lib.require_device()
    _lib.check(lib.bh_step_factor_f64(_mode(mode), _lib.ptr(r), r.shape[0], _lib.ptr(out),
                                      _lib.stream_handle()), "step_factor")
    return _dev.like_input(out.reshape(shape), ratio)


def _fixed_steps(rhs, state: Array, h: Array, n_steps: int, mode, launch, what: str) -> Array:
    g = _resolve(rhs)
    s = _dev.as_device_f64(state, (8,), "state")
    n = s.shape[0]
    hv = _dev.as_device_f64(h, None, "h").reshape(-1)
    if hv.shape[0] != n:
        raise ValueError("h must have shape (n,)")
    if n_steps < 0:
        raise ValueError("n_steps must be >= 0")
    lib = _lib.require_device()
    stream = _lib.stream_handle()
    ld = max(n, 1)
    soa = torch.empty((8, ld), dtype=torch.float64, device=s.device)
    _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(s), _lib.ptr(soa), ld, n, stream), "aos_to_soa")
    _lib.check(launch(lib, g.spacetime, _mode(mode or g.mode), soa, ld, hv, n, int(n_steps),
                      stream), what)
    out = torch.empty_like(s)
    _lib.check(lib.bh_soa_to_aos_f64(_lib.ptr(soa), ld, _lib.ptr(out), n, stream), "soa_to_aos")
    return _dev.like_input(out, state)


def rk4_steps(rhs, state: Array, h: Array, n_steps: int, *, mode: str | None = None) -> Array:
    """``n_steps`` classical RK4 steps (integrators.py:57-73) per ray in one
    kernel launch, the state kept in registers between steps; equals
    applying ``rk4_step`` n_steps times (bitwise in the precise build)."""
    def launch(lib, st, m, soa, ld, hv, n, k, stream):
        return lib.bh_rk4_f64(st.mass, st.spin, m, _lib.ptr(soa), ld, _lib.ptr(hv), n, k, stream)

    return _fixed_steps(rhs, state, h, n_steps, mode, launch, "rk4")


def implicit_midpoint_steps(rhs, state: Array, h: Array, n_steps: int, iterations: int = 4, *,
                            mode: str | None = None) -> Array:
    """``n_steps`` implicit midpoint steps (integrators.py:177-207) per ray
    in one launch; equals applying ``implicit_midpoint_step`` n_steps times
    (bitwise in the precise build)."""
    if iterations < 0:
        raise ValueError("iterations must be >= 0")

    def launch(lib, st, m, soa, ld, hv, n, k, stream):
        return lib.bh_midpoint_f64(st.mass, st.spin, m, _lib.ptr(soa), ld, _lib.ptr(hv), n, k,
                                   int(iterations), stream)

    return _fixed_steps(rhs, state, h, n_steps, mode, launch, "implicit_midpoint")


def implicit_midpoint_step(rhs, state: Array, h: Array, iterations: int = 4) -> Array:
    """One implicit midpoint step, symplectic for Hamiltonian flows: solves
    y1 = y0 + h f((y0 + y1) / 2) by ``iterations`` fixed-point sweeps from an
    explicit half step (integrators.py:177-207)."""
    return implicit_midpoint_steps(rhs, state, h, 1, iterations)


def rk4_step(rhs, state: Array, h: Array) -> Array:
    """One classical RK4 step, (n, 8) -> (n, 8) (integrators.py:57-73)."""
    return rk4_steps(rhs, state, h, 1)


def integrate_dense(rhs, state0: Array, lambda_total: float, rtol: float = 1e-10,
                    atol: float = 1e-12, initial_step: float = 0.05, max_step: float = 2.0,
                    *, mode: str | None = None, chunk: int = 4096) -> list:
    """Adaptively integrate a batch over an affine-parameter interval
    (integrators.py:127-174), on the device.

    Same semantics as the reference: every sample shares the step sequence
    (the largest error ratio of the batch controls it), h is clipped to
    min(h, max_step, lambda_total - elapsed) before each attempt, and the
    result is the list of accepted states starting with a copy of
    ``state0``.  Raises RuntimeError("integrate_dense: step size underflow")
    when the step falls below 1e-14.  ``rhs`` is ``GeodesicRHS(spacetime)``
    or a ``KerrSpacetime``; its arithmetic build is ``mode`` if given, else
    the GeodesicRHS's (precise by default).  The kernel appends up to
    ``chunk`` states per launch and resumes from its saved controller state.
    """
    g = _resolve(rhs)
    s0 = _dev.as_device_f64(state0, (8,), "state0")
    n = s0.shape[0]
    chunks = [s0.clone()[None]]
    if not float(lambda_total) > 0.0:
        return _as_list(chunks, state0)
    if n == 0:
        # the reference's batch maximum of an empty ratio array raises
        raise ValueError("integrate_dense needs at least one state")
    lib = _lib.require_device()
    stream = _lib.stream_handle()
    dev = s0.device
    y = torch.empty((8, n), dtype=torch.float64, device=dev)
    _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(s0), _lib.ptr(y), n, n, stream), "aos_to_soa")
    work = torch.empty(24 * n + 8192, dtype=torch.float64, device=dev)
    traj = torch.empty((chunk, 8, n), dtype=torch.float64, device=dev)
    ctl = torch.tensor([0.0, float(initial_step)], dtype=torch.float64, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    m = _mode(mode or g.mode)
    while True:
        _lib.check(lib.bh_integrate_dense_f64(
            g.spacetime.mass, g.spacetime.spin, m, float(rtol), float(atol), float(max_step),
            float(lambda_total), _lib.ptr(y), n, n, _lib.ptr(work), _lib.ptr(traj), chunk,
            _lib.ptr(ctl), _lib.ptr(count), _lib.ptr(flag), stream), "integrate_dense")
        k = int(count.item())
        f = int(flag.item())
        if k:
            chunks.append(traj[:k].transpose(1, 2).contiguous())
        if f == _lib.DENSE_UNDERFLOW:
            raise RuntimeError("integrate_dense: step size underflow")
        if f == _lib.DENSE_DONE:
            break
    return _as_list(chunks, state0)


def _as_list(chunks, ref) -> list:
    """Concatenated (k, n, 8) chunks as a list of k (n, 8) states, in the
    caller's array type (one device-to-host copy for NumPy callers)."""
    allst = _dev.like_input(torch.cat(chunks), ref)
    return [allst[i] for i in range(allst.shape[0])]

