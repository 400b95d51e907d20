"""The real-time shader's tracing algorithm on MI355X (reference
realtime/reference.py and realtime/settings.py).

The reference's GLSL fragment shader integrates with fixed-order RK4 and a
distance-based step heuristic; ``realtime/reference.py`` mirrors it in
NumPy so its accuracy can be checked against the adaptive tracer.  Here the
same algorithm runs per ray in ``bh_trace_shader_f64`` -- a path for
offline video with the real-time look, at the real-time presets' cost.
Only the settings the tracing reads are kept (no window, UI or shading).
"""

from __future__ import annotations

from dataclasses import dataclass, replace
from enum import Enum

import numpy
import torch

from . import _dev, _lib
from .backend import Array
from .kerr import KerrSpacetime, _mode
from .tracer import _order_tensor

MAX_STEPS_HARD_LIMIT = 4096  # settings.py:52-54


class QualityPreset(str, Enum):
    """settings.py:24-30."""

    LOW = "low"
    MEDIUM = "medium"
    HIGH = "high"
    ULTRA = "ultra"


_PRESET_VALUES = {  # settings.py:33-50
    QualityPreset.LOW: {"max_steps": 256, "step_scale": 0.35, "resolution_scale": 0.5},
    QualityPreset.MEDIUM: {"max_steps": 512, "step_scale": 0.22, "resolution_scale": 0.75},
    QualityPreset.HIGH: {"max_steps": 1024, "step_scale": 0.14, "resolution_scale": 1.0},
    QualityPreset.ULTRA: {"max_steps": 2048, "step_scale": 0.08, "resolution_scale": 1.0},
}


@dataclass
class RenderSettings:
    """The tracing-relevant tunables of the real-time tracer (settings.py:58-173)."""

    spin: float = 0.9
    fov_degrees: float = 60.0
    max_steps: int = 512
    step_scale: float = 0.22
    min_step: float = 1e-4
    max_step: float = 4.0
    resolution_scale: float = 0.75
    capture_margin: float = 2e-2
    escape_radius: float = 0.0
    momentum_bailout: float = 1e3
    interior_stop: float = 0.02
    disk_enabled: bool = True
    disk_outer_radius: float = 18.0

    def validate(self) -> None:
        """The range checks of settings.py:135-163 that concern tracing."""
        if not -1.0 <= self.spin <= 1.0:
            raise ValueError("spin must be in [-1, 1]")
        if not 1 <= self.max_steps <= MAX_STEPS_HARD_LIMIT:
            raise ValueError(f"max_steps must be in [1, {MAX_STEPS_HARD_LIMIT}]")
        if self.step_scale <= 0.0:
            raise ValueError("step_scale must be positive")
        if not 0.0 < self.min_step <= self.max_step:
            raise ValueError("require 0 < min_step <= max_step")
        if self.capture_margin < 0.0:
            raise ValueError("capture_margin must be non-negative")
        if self.momentum_bailout <= 1.0:
            raise ValueError("momentum_bailout must exceed 1")
        if not 1e-4 <= self.interior_stop <= 1.0:
            raise ValueError("interior_stop must be in [1e-4, 1] M")

    def apply_preset(self, preset: QualityPreset) -> "RenderSettings":
        return replace(self, **_PRESET_VALUES[preset])

    def effective_escape_radius(self, camera_distance: float) -> float:
        if self.escape_radius > 0.0:
            return self.escape_radius
        return max(2.0 * camera_distance, 100.0)


@dataclass(frozen=True)
class ReferenceResult:
    """Status, final positions and momenta, and disk hits (crossing radius,
    linearly interpolated position and momentum; NaN for non-hits) when a
    disk is configured (reference.py:23-38)."""

    status: Array
    positions: Array
    momenta: Array
    hit_radii: Array | None = None
    hit_positions: Array | None = None
    hit_momenta: Array | None = None


def trace_like_shader(spacetime: KerrSpacetime, state0: Array, settings: RenderSettings,
                      escape_radius: float, disk_radii: tuple[float, float] | None = None,
                      interior_stop: float | None = None, *, mode: str = "precise",
                      ray_order: Array | None = None) -> ReferenceResult:
    """Trace rays exactly as the real-time fragment shader does
    (reference.py:41-188); see bh_trace_shader_f64.  ``mode`` "precise"
    rounds as the reference's NumPy mirror, "fast" contracts to FMAs."""
    s0 = _dev.as_device_f64(state0, (8,), "state0")
    n = s0.shape[0]
    dev = s0.device
    if interior_stop is None:
        capture_radius = spacetime.outer_horizon_radius * (1.0 + settings.capture_margin)
        max_steps = settings.max_steps
    else:
        capture_radius = float(interior_stop)
        max_steps = min(settings.max_steps, 1024)  # reference.py:101-103
    flags = (_lib.TILE_DISK if disk_radii is not None else 0) | (
        _lib.TILE_INTERIOR if interior_stop is not None else 0)
    inner, outer = disk_radii if disk_radii is not None else (0.0, 0.0)
    p = _lib.ShaderParams(spacetime.mass, spacetime.spin, capture_radius, float(escape_radius),
                          float(settings.step_scale), float(settings.min_step),
                          float(settings.max_step), float(settings.momentum_bailout),
                          float(inner), float(outer), int(max_steps), flags, _mode(mode))
    lib = _lib.require_device()
    stream = _lib.stream_handle()
    ld = max(n, 1)
    soa = torch.empty((8, ld), dtype=torch.float64, device=dev)
    _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(s0), _lib.ptr(soa), ld, n, stream), "aos_to_soa")
    out = torch.empty_like(soa)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    hit = torch.empty((n, 8), dtype=torch.float64, device=dev) if disk_radii is not None else None
    _lib.check(lib.bh_trace_shader_f64(p, _lib.ptr(soa), ld, _lib.ptr(out), ld, _lib.ptr(status),
                                       _lib.ptr(hit), _lib.ptr(_order_tensor(ray_order, n, dev)),
                                       n, stream), "trace_like_shader")
    states = torch.empty((n, 8), dtype=torch.float64, device=dev)
    _lib.check(lib.bh_soa_to_aos_f64(_lib.ptr(out), ld, _lib.ptr(states), n, stream),
               "soa_to_aos")
    conv = lambda t: _dev.like_input(t, state0)  # noqa: E731
    return ReferenceResult(
        status=conv(status), positions=conv(states[:, 1:4].contiguous()),
        momenta=conv(states[:, 4:8].contiguous()),
        hit_radii=None if hit is None else conv(hit[:, 0].contiguous()),
        hit_positions=None if hit is None else conv(hit[:, 1:4].contiguous()),
        hit_momenta=None if hit is None else conv(hit[:, 4:8].contiguous()))
