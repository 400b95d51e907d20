"""Many trace jobs in one persistent launch (``bh_trace_batched_f64``).

The reference traces one job per ``trace_rays`` call (tracer.py:50-157); a
workload of many small or uneven jobs (different spins, cameras, tolerances,
ray counts) then runs job after job, and each call's tail -- the few rays
still orbiting near the photon shell -- leaves the GPU mostly idle.  Here the
rays of all jobs form one queue: a persistent grid refills wavefront lanes
from it as rays retire, and every ray is still traced by exactly the per-ray
controller of trace_rays, so each job's results equal what ``trace_rays``
returns for that job alone.

``BatchPlan`` uploads everything once and ``launch()`` only enqueues work,
so it can be timed or captured; ``trace_batch`` is the one-call form.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy
import torch

from . import _dev, _lib
from .backend import Array
from .kerr import KerrSpacetime
from .tracer import TraceResult, Tune, check_max_steps, make_params, resolve_mode

MAX_JOBS_PER_LAUNCH = 256


@dataclass
class TraceJob:
    """One trace_rays call's arguments (tracer.py:50-61), plus an optional
    affine-parameter end (see tracer.integrate_rays)."""

    spacetime: KerrSpacetime
    state0: Array
    escape_radius: float
    max_steps: int = 20000
    rtol: float = 1e-9
    atol: float = 1e-12
    initial_step: float = 0.1
    max_step: float = 2.0
    min_step: float = 1e-10
    capture_margin: float = 1e-3
    lambda_end: float = math.inf


def _initial_radii(state0) -> numpy.ndarray:
    """|x| of every initial position, on the host."""
    pos = state0[:, 1:4]
    if isinstance(pos, torch.Tensor):
        pos = pos.detach().to("cpu")
    return numpy.linalg.norm(numpy.asarray(pos, dtype=numpy.float64), axis=1)


def expected_steps(job: TraceJob) -> numpy.ndarray | float:
    """Expected accepted steps of a job's rays, for ranking only.  A ray
    running to escape covers its distance at the largest step,
    ``escape_radius / max_step``, plus the steps near the hole, which a
    fifth-order controller's step scaling makes grow as rtol^-1/5: one value
    for the job, ``escape_radius / max_step + 6 rtol^-1/5``.  A ray
    integrated to an affine end takes ``lambda_end / h`` steps, where the
    controller's step on a bound orbit of radius r is about one dynamical
    time r^3/2 scaled by rtol^1/5, capped at max_step: one value per ray,
    from its initial radius."""
    if math.isfinite(job.lambda_end):
        r = numpy.maximum(_initial_radii(job.state0), 1e-3)
        h = numpy.minimum(job.max_step, r ** 1.5 * job.rtol ** 0.2)
        return job.lambda_end / h
    return job.escape_radius / job.max_step + 6.0 * job.rtol ** -0.2


def longest_first(jobs: list[TraceJob], offsets) -> numpy.ndarray:
    """Queue order handing out the rays with the longest expected traces
    first (the longest-processing-time rule, ``expected_steps``), so that a
    launch does not end on a long trace that started late: jobs traced to
    escape are ranked as a whole, rays integrated to an affine end (timelike
    orbits) one by one, innermost first; ties keep the given order.  No
    result depends on the order."""
    n = int(offsets[-1]) if len(offsets) else 0
    est = numpy.empty(n, dtype=numpy.float64)
    for i, j in enumerate(jobs):
        lo, hi = int(offsets[i]), int(offsets[i + 1])
        if hi > lo:
            est[lo:hi] = expected_steps(j)
    return numpy.argsort(-est, kind="stable").astype(numpy.int64)


class _Group:
    """Jobs of one arithmetic mode, traced by one persistent launch."""

    def __init__(self, jobs, idx, mode, dev, order_fn):
        self.idx = idx
        self.mode = mode
        for i in idx:
            check_max_steps(jobs[i].max_steps)
        states = [_dev.as_device_f64(jobs[i].state0, (8,), "state0") for i in idx]
        counts = [s.shape[0] for s in states]
        self.offsets_host = numpy.concatenate([[0], numpy.cumsum(counts)]).astype(numpy.int64)
        n = int(self.offsets_host[-1])
        self.n = n
        lib = _lib.require_device()
        stream = _lib.stream_handle()
        ld = max(n, 1)
        self.soa_in = torch.empty((8, ld), dtype=torch.float64, device=dev)
        params = []
        for k, i in enumerate(idx):
            j = jobs[i]
            p = make_params(j.spacetime, j.escape_radius, j.max_steps, j.rtol, j.atol,
                            j.initial_step, j.max_step, j.min_step, j.capture_margin, mode,
                            j.lambda_end)
            params.append(p)
            lo = int(self.offsets_host[k])
            if counts[k]:
                view = self.soa_in[:, lo:]
                _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(states[k]), _lib.ptr(view), ld,
                                                 counts[k], stream), "aos_to_soa")
        # initial-radius check per job (tracer.py:94-99)
        cnt = torch.zeros(len(idx), dtype=torch.int32, device=dev)
        for k, p in enumerate(params):
            lo = int(self.offsets_host[k])
            if counts[k]:
                _lib.check(lib.bh_count_outside_f64(p, _lib.ptr(self.soa_in[:, lo:]), ld,
                                                    counts[k], _lib.ptr(cnt[k:k + 1]), stream),
                           "count_outside")
        if counts and int(cnt.sum().item()) > 0:
            raise ValueError("escape_radius must exceed all initial radii")
        arr = (_lib.TraceParams * len(params))(*params)
        raw = numpy.frombuffer(bytes(arr), dtype=numpy.uint8).copy()
        self.jobs_dev = torch.from_numpy(raw).to(dev)
        self.offsets_dev = torch.from_numpy(self.offsets_host).to(dev)
        self.soa_out = torch.empty_like(self.soa_in)
        self.status = torch.empty(n, dtype=torch.int32, device=dev)
        self.steps = torch.empty(n, dtype=torch.int64, device=dev)
        self.attempts = torch.empty(n, dtype=torch.int64, device=dev)
        self.iters = torch.zeros(len(idx), dtype=torch.int64, device=dev)
        self.workspace = torch.zeros(1, dtype=torch.int64, device=dev)
        self.order = None
        if order_fn is not None and n:
            o = numpy.asarray(order_fn([jobs[i] for i in idx], self.offsets_host), numpy.int64)
            if o.shape != (n,) or not numpy.array_equal(numpy.sort(o), numpy.arange(n)):
                raise ValueError("order must be a permutation of the batch's rays")
            self.order = torch.from_numpy(o).to(dev)

    def launch(self, tune: _lib.TuneParams, stream: int) -> None:
        lib = _lib.load()
        if tune.lanes_per_ray > 1 and self.mode != "precise":
            # the quad tracer exists in the precise build only
            tune = _lib.TuneParams(tune.block_size, tune.rays_per_lane, tune.refill_min, 1)
        _lib.check(lib.bh_trace_batched_f64(
            _lib.ptr(self.jobs_dev), _lib.ptr(self.offsets_dev), len(self.idx),
            _lib.MODE_FAST if self.mode == "fast" else _lib.MODE_PRECISE, tune,
            _lib.ptr(self.soa_in), self.soa_in.shape[1], _lib.ptr(self.soa_out),
            self.soa_out.shape[1], _lib.ptr(self.status), _lib.ptr(self.steps),
            _lib.ptr(self.attempts), _lib.ptr(self.iters), _lib.ptr(self.order),
            _lib.ptr(self.workspace), self.n, stream), "trace_batch")


class BatchPlan:
    """Device-resident inputs, job tables and outputs of a batched trace.

    Jobs are grouped by resolved arithmetic mode ("auto" resolves per job by
    its rtol); each group of up to 256 jobs is one persistent launch.  With
    ``concurrent=True`` the groups run on separate streams.  ``order_fn(jobs,
    offsets)`` returns the order in which a group's rays are handed out (a
    permutation of its rays; default ``longest_first``, None for the given
    order); it changes no result.
    """

    def __init__(self, jobs: list[TraceJob], mode: str = "auto", tune: Tune | None = None,
                 order_fn=longest_first, concurrent: bool = True):
        if not jobs:
            raise ValueError("no jobs")
        dev = _dev.device()
        self.jobs = jobs
        self.tune = tune or Tune()
        by_mode: dict[str, list[int]] = {}
        for i, j in enumerate(jobs):
            by_mode.setdefault(resolve_mode(mode, j.rtol), []).append(i)
        self.groups = []
        for m, idx in sorted(by_mode.items()):
            for c in range(0, len(idx), MAX_JOBS_PER_LAUNCH):
                self.groups.append(_Group(jobs, idx[c:c + MAX_JOBS_PER_LAUNCH], m, dev, order_fn))
        self.streams = None
        if concurrent and len(self.groups) > 1:
            self.streams = [torch.cuda.Stream(device=dev) for _ in self.groups]

    def launch(self, tune: Tune | None = None) -> None:
        """Enqueue every group (no host synchronisation); ``tune`` overrides
        the plan's launch shape (results do not depend on it)."""
        t = (tune or self.tune).as_c()
        cur = torch.cuda.current_stream()
        if self.streams is None:
            for g in self.groups:
                g.launch(t, _lib.stream_handle(cur))
            return
        ev = torch.cuda.Event()
        ev.record(cur)
        for g, s in zip(self.groups, self.streams):
            s.wait_event(ev)
            g.launch(t, _lib.stream_handle(s))
        for s in self.streams:
            cur.wait_stream(s)

    def results(self, as_numpy: bool | None = None) -> list[TraceResult]:
        """Per-job TraceResults in the order the jobs were given."""
        out: list[TraceResult | None] = [None] * len(self.jobs)
        lib = _lib.load()
        stream = _lib.stream_handle()
        for g in self.groups:
            iters = _dev.to_numpy(g.iters)
            for k, i in enumerate(g.idx):
                lo, hi = int(g.offsets_host[k]), int(g.offsets_host[k + 1])
                n = hi - lo
                states = torch.empty((n, 8), dtype=torch.float64, device=g.soa_in.device)
                if n:
                    _lib.check(lib.bh_soa_to_aos_f64(_lib.ptr(g.soa_out[:, lo:]),
                                                     g.soa_out.shape[1], _lib.ptr(states), n,
                                                     stream), "soa_to_aos")
                ref = self.jobs[i].state0
                conv = _dev.to_numpy if (
                    as_numpy or (as_numpy is None and not isinstance(ref, torch.Tensor))) \
                    else (lambda t: t)
                out[i] = TraceResult(states=conv(states), status=conv(g.status[lo:hi]),
                                     steps=conv(g.steps[lo:hi]), iterations=int(iters[k]),
                                     attempts=conv(g.attempts[lo:hi]))
        return out  # type: ignore[return-value]

    @property
    def total_rays(self) -> int:
        return sum(g.n for g in self.groups)


def trace_batch(jobs: list[TraceJob], *, mode: str = "auto", tune: Tune | None = None,
                order_fn=longest_first) -> list[TraceResult]:
    """Trace every job to termination in persistent batched launches; each
    job's TraceResult equals trace_rays on that job alone."""
    plan = BatchPlan(jobs, mode=mode, tune=tune, order_fn=order_fn)
    plan.launch()
    return plan.results()
