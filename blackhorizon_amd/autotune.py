"""Launch-shape autotuner for the plain and persistent batched tracers.

The tracer's results never depend on its launch shape (every ray runs the
same per-ray controller); only the speed does.  The shape has three knobs
(``tracer.Tune``):

* ``block_size`` -- threads per block, a multiple of 64 up to 256 (the
  kernels' launch bound);
* ``rays_per_lane`` (plain kernel) -- rays of a wavefront's range handed to
  its lanes in turn: 1 = one ray per lane, the hardware dispatcher refills;
  >1 = in-wave refill from the wave's range;
* ``refill_min`` -- idle lanes a wavefront waits for before refilling (plain
  kernel with rays_per_lane > 1, and the persistent kernel's global queue);
* ``lanes_per_ray`` (batched kernel, precise build) -- 1, or 4 for the quad
  tracer (one DPP quad of lanes per ray), which shortens each ray's
  attempts at the cost of throughput: it pays when a few long traces set
  the launch's time.

``autotune(plan)`` times every candidate shape on a prepared plan
(``tracer.TracePlan`` or ``batch.BatchPlan``: device-resident inputs, so only
the kernels are timed, with HIP events on the launch stream), returns the
fastest and records the table in a JSON cache keyed by device, kernel, mode
and a caller-supplied workload key, so that later runs reuse the choice
without timing.
"""

from __future__ import annotations

import json
import os
import statistics
from dataclasses import asdict, dataclass, field

import torch

from . import _lib
from .tracer import Tune

BLOCKS = (64, 128, 192, 256)
PLAIN_RAYS_PER_LANE = (1, 2, 4)
REFILL_MINS = (1, 16, 64)
BATCH_REFILL_MINS = (1, 4, 16, 32, 64)

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "blackhorizon_amd",
                             "autotune.json")


def plain_candidates() -> list[Tune]:
    """Shapes for the plain kernel; the default (256, 1, 1) comes first so
    that it wins ties."""
    out = [Tune()]
    for b in BLOCKS:
        for rpl in PLAIN_RAYS_PER_LANE:
            for rm in (REFILL_MINS if rpl > 1 else (1,)):
                t = Tune(b, rpl, rm)
                if t not in out:
                    out.append(t)
    return out


QUAD_REFILL_MINS = (1, 4, 16)


def batch_candidates(quad: bool = False) -> list[Tune]:
    """Shapes for the persistent batched kernel (rays_per_lane unused); with
    ``quad`` (a plan with a precise group) also the quad tracer's."""
    out = [Tune()]
    for b in BLOCKS:
        for rm in BATCH_REFILL_MINS:
            t = Tune(b, 1, rm)
            if t not in out:
                out.append(t)
    if quad:
        out += [Tune(256, 1, rm, 4) for rm in QUAD_REFILL_MINS]
    return out


@dataclass
class TuneResult:
    """The chosen shape, its time and the whole table (ms per launch)."""

    key: str
    best: Tune
    best_ms: float
    table: list = field(default_factory=list)  # [(Tune as dict, ms), ...]
    cached: bool = False

    def as_dict(self) -> dict:
        return dict(key=self.key, best=asdict(self.best), best_ms=self.best_ms,
                    table=self.table, cached=self.cached)


def _plan_kind(plan) -> str:
    return "batched" if hasattr(plan, "groups") else "plain"


def _plan_mode(plan) -> str:
    if hasattr(plan, "groups"):
        return "+".join(sorted({g.mode for g in plan.groups}))
    return plan.mode


def cache_key(plan, workload_key: str) -> str:
    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(props, "gcnArchName", "") or props.name
    return "|".join([arch, str(props.multi_processor_count), _plan_kind(plan),
                     _plan_mode(plan), workload_key])


def select(table: list[tuple[Tune, float]]) -> tuple[Tune, float]:
    """The fastest entry; the earliest one on ties."""
    if not table:
        raise ValueError("empty autotune table")
    best = min(range(len(table)), key=lambda i: (table[i][1], i))
    return table[best]


def time_candidates(plan, candidates: list[Tune], reps: int = 3, warmup: int = 1,
                    timer=None) -> list[tuple[Tune, float]]:
    """Median ms per launch of every candidate (HIP events on the current
    stream; ``timer(plan, tune, reps, warmup)`` replaces the measurement)."""
    timer = timer or _event_timer
    return [(t, timer(plan, t, reps, warmup)) for t in candidates]


def _event_timer(plan, tune: Tune, reps: int, warmup: int) -> float:
    cur = torch.cuda.current_stream()
    for _ in range(warmup):
        plan.launch(tune)
    ms = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(cur)
        plan.launch(tune)
        e1.record(cur)
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    return statistics.median(ms)


def load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_cache(path: str, key: str, result: TuneResult) -> None:
    data = load_cache(path)
    data[key] = result.as_dict()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(plan, workload_key: str = "default", cache: str | None = DEFAULT_CACHE,
             candidates: list[Tune] | None = None, reps: int = 3, warmup: int = 1,
             retune: bool = False, timer=None) -> TuneResult:
    """Pick the fastest launch shape for ``plan`` (a TracePlan or BatchPlan).

    With a ``cache`` path, a stored choice for the same device, kernel, mode
    and ``workload_key`` is returned without timing unless ``retune``.  The
    plan's outputs hold the last timed launch's results, which equal an
    untuned launch's.  The plan's own shape is set to the choice.
    """
    _lib.require_device()
    key = cache_key(plan, workload_key)
    if cache and not retune:
        hit = load_cache(cache).get(key)
        if hit is not None:
            best = Tune(**hit["best"])
            plan.tune = best
            return TuneResult(key, best, float(hit["best_ms"]), hit.get("table", []), True)
    if candidates is None:
        if _plan_kind(plan) == "plain":
            candidates = plain_candidates()
        else:
            candidates = batch_candidates(quad=any(g.mode == "precise" for g in plan.groups))
    table = time_candidates(plan, candidates, reps, warmup, timer)
    best, best_ms = select(table)
    res = TuneResult(key, best, best_ms, [(asdict(t), ms) for t, ms in table])
    plan.tune = best
    if cache:
        store_cache(cache, key, res)
    return res
