"""Batched adaptive geodesic tracing on MI355X (reference tracer.py).

``trace_rays`` keeps the reference signature and result type
(tracer.py:50-157).  The whole controller loop -- DP5(4) attempt, error
ratio, step-size update, capture/escape/failure classification -- runs per
ray inside one HIP kernel (``bh_trace_plain_f64``); the host only validates,
moves buffers and reads back ``iterations``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from enum import IntEnum

import torch

from . import _dev, _lib
from .backend import Array
from .kerr import KerrSpacetime, _mode


class RayStatus(IntEnum):
    """Terminal state of a traced geodesic (tracer.py:21-30)."""

    IN_FLIGHT = 0
    CAPTURED = 1
    ESCAPED = 2
    MAX_STEPS = 3
    FAILED = 4
    DISK = 5
    TERMINATED = 6


@dataclass(frozen=True)
class TraceResult:
    """tracer.py:33-47, plus ``attempts`` (accepted + rejected steps per ray)."""

    states: Array
    status: Array
    steps: Array
    iterations: int
    attempts: Array | None = None


@dataclass(frozen=True)
class Tune:
    """Launch shape of the tracer kernels: threads per block, rays each lane
    of the plain tracer takes in turn, idle lanes (quads) a wavefront
    collects before refilling, and lanes per ray of the batched tracer (4:
    a DPP quad per ray, precise build only).  No result depends on it."""

    block_size: int = 256
    rays_per_lane: int = 1
    refill_min: int = 1
    lanes_per_ray: int = 1

    def as_c(self) -> _lib.TuneParams:
        return _lib.TuneParams(self.block_size, self.rays_per_lane, self.refill_min,
                               self.lanes_per_ray)


# Below this rtol the fast build's last-ulp RHS rounding, amplified by the
# embedded error estimate (~1e-16 / rtol) into the step sizes, can move final
# states by more than 1e-7 relative to the reference (observed at rtol 1e-9
# with atol 1e-12); "auto" then uses the precise build.
FAST_MIN_RTOL = 1e-8


def resolve_mode(mode: str, rtol: float) -> str:
    if mode == "auto":
        return "fast" if rtol >= FAST_MIN_RTOL else "precise"
    if mode not in ("fast", "precise"):
        raise ValueError("mode must be 'auto', 'fast' or 'precise'")
    return mode


MAX_STEPS_ABS = 1 << 60


def check_max_steps(max_steps) -> int:
    """max_steps as an int, refused (ValueError) outside [-2^60, 2^60]: the
    kernels cap attempts at 4 * max_steps in int64 (tracer.py:101).  The C
    ABI refuses such a value for one job; a batched launch reads its job
    table on the device, so the host checks every job."""
    max_steps = int(max_steps)
    if not -MAX_STEPS_ABS <= max_steps <= MAX_STEPS_ABS:
        raise ValueError("max_steps must lie in [-2**60, 2**60]")
    return max_steps


def make_params(spacetime: KerrSpacetime, escape_radius: float, max_steps: int, rtol: float,
                atol: float, initial_step: float, max_step: float, min_step: float,
                capture_margin: float, mode: str = "auto",
                lambda_end: float = math.inf) -> _lib.TraceParams:
    max_steps = check_max_steps(max_steps)
    # tracer.py:84 computes the capture radius in host double precision
    capture_radius = spacetime.outer_horizon_radius * (1.0 + capture_margin)
    return _lib.TraceParams(spacetime.mass, spacetime.spin, float(escape_radius),
                            capture_radius, float(rtol), float(atol), float(initial_step),
                            float(max_step), float(min_step), int(max_steps),
                            _mode(resolve_mode(mode, rtol)), 0, float(lambda_end))


def _order_tensor(ray_order, n: int, dev):
    """Validated int64 device permutation of [0, n), or None."""
    if ray_order is None:
        return None
    o = torch.as_tensor(ray_order, dtype=torch.int64).to(dev).reshape(-1).contiguous()
    if o.shape[0] != n:
        raise ValueError("ray_order must have one entry per ray")
    if n and not torch.equal(torch.sort(o).values, torch.arange(n, device=dev)):
        raise ValueError("ray_order must be a permutation of range(n)")
    return o


class TracePlan:
    """Device-resident inputs, outputs and parameters of one plain trace.

    The constructor uploads ``state0`` (SoA), checks the initial radii
    (tracer.py:94-99) and allocates the outputs; ``launch()`` only enqueues
    the tracer kernel on the current stream -- no allocation, no host
    synchronisation, no memset -- so it can be repeated, timed or captured in
    a HIP graph (``graph.TraceGraph``); ``result()`` converts the outputs to
    a TraceResult.  Arguments are trace_rays'.
    """

    def __init__(self, spacetime: KerrSpacetime, state0: Array, escape_radius: float,
                 max_steps: int = 20000, rtol: float = 1e-9, atol: float = 1e-12,
                 initial_step: float = 0.1, max_step: float = 2.0, min_step: float = 1e-10,
                 capture_margin: float = 1e-3, *, mode: str = "auto", tune: Tune | None = None,
                 ray_order: Array | None = None, lambda_end: float = math.inf):
        self.params = make_params(spacetime, escape_radius, max_steps, rtol, atol, initial_step,
                                  max_step, min_step, capture_margin, mode, lambda_end)
        self.mode = resolve_mode(mode, rtol)
        self.state0 = state0
        s0 = _dev.as_device_f64(state0, (8,), "state0")
        n = s0.shape[0]
        self.n = n
        dev = s0.device
        lib = _lib.require_device()
        stream = _lib.stream_handle()
        self.order = _order_tensor(ray_order, n, dev)
        ld = max(n, 1)
        self.soa_in = torch.empty((8, ld), dtype=torch.float64, device=dev)
        _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(s0), _lib.ptr(self.soa_in), ld, n, stream),
                   "aos_to_soa")
        count = torch.zeros(1, dtype=torch.int32, device=dev)
        _lib.check(lib.bh_count_outside_f64(self.params, _lib.ptr(self.soa_in), ld, n,
                                            _lib.ptr(count), stream), "count_outside")
        if int(count.item()) > 0:
            raise ValueError("escape_radius must exceed all initial radii")
        self.soa_out = torch.empty_like(self.soa_in)
        self.status = torch.empty(n, dtype=torch.int32, device=dev)
        self.steps = torch.empty(n, dtype=torch.int64, device=dev)
        self.attempts = torch.empty(n, dtype=torch.int64, device=dev)
        self.iters = torch.zeros(1, dtype=torch.int64, device=dev)
        self.tune = tune or Tune()

    def launch(self, tune: Tune | None = None) -> None:
        """Enqueue the trace on the current stream (``tune`` overrides the
        plan's launch shape; results do not depend on it)."""
        lib = _lib.load()
        ld = self.soa_in.shape[1]
        _lib.check(lib.bh_trace_plain_f64(
            self.params, (tune or self.tune).as_c(), _lib.ptr(self.soa_in), ld,
            _lib.ptr(self.soa_out), ld, _lib.ptr(self.status), _lib.ptr(self.steps),
            _lib.ptr(self.attempts), _lib.ptr(self.iters), _lib.ptr(self.order), self.n,
            _lib.stream_handle()), "trace_rays")

    def result(self) -> TraceResult:
        """The last launch's outputs, NumPy or device tensors like state0."""
        lib = _lib.load()
        n = self.n
        states = torch.empty((n, 8), dtype=torch.float64, device=self.soa_out.device)
        _lib.check(lib.bh_soa_to_aos_f64(_lib.ptr(self.soa_out), self.soa_out.shape[1],
                                         _lib.ptr(states), n, _lib.stream_handle()),
                   "soa_to_aos")
        ref = self.state0
        return TraceResult(states=_dev.like_input(states, ref),
                           status=_dev.like_input(self.status.clone(), ref),
                           steps=_dev.like_input(self.steps.clone(), ref),
                           iterations=int(self.iters.item()),
                           attempts=_dev.like_input(self.attempts.clone(), ref))


def trace_rays(
    spacetime: KerrSpacetime,
    state0: Array,
    escape_radius: float,
    max_steps: int = 20000,
    rtol: float = 1e-9,
    atol: float = 1e-12,
    initial_step: float = 0.1,
    max_step: float = 2.0,
    min_step: float = 1e-10,
    capture_margin: float = 1e-3,
    *,
    mode: str = "auto",
    tune: Tune | None = None,
    ray_order: Array | None = None,
    lambda_end: float = math.inf,
) -> TraceResult:
    """Trace a batch of geodesics to termination (tracer.py:50-157).

    ``state0`` is an (n, 8) NumPy array (copied to the device; results come
    back as NumPy) or a float64 device tensor (results stay on the device).
    It is never modified.  Raises ValueError if any initial Kerr-Schild
    radius is >= ``escape_radius`` (tracer.py:94-99).  ``mode``: "precise"
    rounds every operation as the reference's NumPy expressions do; "fast"
    contracts to FMAs and shares reciprocals; "auto" (default) takes fast
    for rtol >= 1e-8 and precise below, where the fast build's rounding
    would move final states by more than 1e-7.  ``ray_order`` (a
    permutation) sets which rays share a wavefront -- e.g.
    ``PinholeCamera.tile_order()`` -- without changing any result.
    ``lambda_end`` (affine parameter) ends rays that reach it with status
    TERMINATED (see integrate_rays).  Without ``tune`` and ``ray_order``, a
    precise-build trace of at most one wavefront per compute unit runs on
    the quad tracer (a DPP quad of lanes per ray, the same results bit for
    bit), which shortens the long rays such a launch waits for.
    """
    if tune is None and ray_order is None and _quads_pay(state0, mode, rtol, 1):
        return _trace_on_quads(spacetime, state0, escape_radius, max_steps, rtol, atol,
                               initial_step, max_step, min_step, capture_margin, lambda_end)
    plan = TracePlan(spacetime, state0, escape_radius, max_steps, rtol, atol, initial_step,
                     max_step, min_step, capture_margin, mode=mode, tune=tune,
                     ray_order=ray_order, lambda_end=lambda_end)
    plan.launch()
    return plan.result()


def _quads_pay(state0, mode: str, rtol: float, waves_per_cu: int) -> bool:
    """A precise-build launch of at most ``waves_per_cu`` one-lane
    wavefronts per compute unit: its time is its longest rays', which the
    quad tracer shortens (one lane per ray wins once the launch fills the
    chip: camera rays at 65,536 rays, where the tail no longer sets it)."""
    if resolve_mode(mode, rtol) != "precise":
        return False
    shape = getattr(state0, "shape", ())
    n = int(shape[0]) if len(shape) == 2 and shape[1] == 8 else 0
    cus = torch.cuda.get_device_properties(_dev.device()).multi_processor_count
    return 0 < n and -(-n // 64) <= waves_per_cu * cus


def _trace_on_quads(spacetime, state0, escape_radius, max_steps, rtol, atol, initial_step,
                    max_step, min_step, capture_margin, lambda_end) -> TraceResult:
    from .batch import TraceJob, trace_batch

    job = TraceJob(spacetime, state0, escape_radius, max_steps, rtol, atol, initial_step,
                   max_step, min_step, capture_margin, lambda_end)
    return trace_batch([job], mode="precise", tune=Tune(lanes_per_ray=4))[0]


def integrate_rays(
    spacetime: KerrSpacetime,
    state0: Array,
    lambda_total: float,
    rtol: float = 1e-10,
    atol: float = 1e-12,
    initial_step: float = 0.05,
    max_step: float = 2.0,
    min_step: float = 1e-14,
    max_steps: int = 10_000_000,
    escape_radius: float = math.inf,
    capture_margin: float = 1e-3,
    *,
    mode: str = "auto",
    tune: Tune | None = None,
) -> TraceResult:
    """Integrate every ray over the affine interval [0, lambda_total] with
    its own adaptive step: per ray, exactly integrate_dense
    (integrators.py:127-174) run on that ray alone -- steps clipped to
    min(h, max_step, lambda_total - lambda), the same error ratio and step
    factor, defaults rtol 1e-10, atol 1e-12, initial_step 0.05 -- except
    that each ray keeps its own step instead of the batch maximum.  Rays
    that reach lambda_total end TERMINATED; a ray whose step falls below
    ``min_step`` (integrate_dense raises at 1e-14) ends FAILED; rays that
    fall through the horizon end CAPTURED (and, with a finite
    ``escape_radius``, rays that leave end ESCAPED).  For timelike rays
    normalised to H = -1/2 the affine parameter is proper time.

    Without ``tune``, a precise-build integration of at most one wavefront
    per SIMD runs on the quad tracer (a DPP quad of lanes per ray, the same
    results bit for bit): such a launch is as long as its longest ray, and
    the quad shortens each of its attempts (65,536 orbits to tau = 1e4:
    208 ms against 289 ms)."""
    if not lambda_total > 0.0:
        raise ValueError("lambda_total must be positive")
    if tune is None and _quads_pay(state0, mode, rtol, 4):
        return _trace_on_quads(spacetime, state0, escape_radius, max_steps, rtol, atol,
                               min(initial_step, max_step), max_step, min_step, capture_margin,
                               float(lambda_total))
    return trace_rays(spacetime, state0, escape_radius, max_steps=max_steps, rtol=rtol,
                      atol=atol, initial_step=min(initial_step, max_step), max_step=max_step,
                      min_step=min_step, capture_margin=capture_margin, mode=mode, tune=tune,
                      lambda_end=float(lambda_total))

