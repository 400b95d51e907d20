"""Offline-renderer tracing on MI355X (reference offline/render.py).

The reference's offline renderer traces its frames with its own copy of
the adaptive DP5(4) loop: classification before each step, equatorial
disk crossings refined by bisection along RK4 sub-steps, flat-space escape
directions, and a second, tighter pass over rays that exhaust their step
budget near the photon shell (render.py:184-349, 592-630).  Here the whole
per-ray loop runs in ``bh_trace_tile_f64``; this module keeps the
renderer's tracing interface (``OfflineSettings``, ``FlyCamera``,
``subpixel_directions``) and returns what the renderer's emission stage
consumes: status, escape directions, disk hit positions and momenta.
Emission (disk spectra, starfield) and image post-processing are not part
of this path.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy
import torch

from . import _dev, _lib
from .backend import Array
from .geodesics import build_state, null_momentum_from_velocity
from .kerr import KerrSpacetime, _mode
from .tracer import RayStatus, Tune, _order_tensor, resolve_mode

_WORLD_UP = numpy.array([0.0, 0.0, 1.0])
_PITCH_LIMIT = math.radians(89.0)


@dataclass
class FlyCamera:
    """Position and yaw/pitch view direction (realtime/fly_camera.py:20-107);
    the orientation math the offline renderer uses."""

    position: numpy.ndarray
    yaw: float = math.pi
    pitch: float = 0.0

    @classmethod
    def from_orbit(cls, distance: float, inclination_degrees: float,
                   azimuth_degrees: float = 0.0) -> "FlyCamera":
        """On the sphere of radius ``distance`` at the given polar angle and
        azimuth, looking at the origin (fly_camera.py:36-62)."""
        inc = math.radians(inclination_degrees)
        az = math.radians(azimuth_degrees)
        pos = numpy.array([distance * math.sin(inc) * math.cos(az),
                           distance * math.sin(inc) * math.sin(az),
                           distance * math.cos(inc)])
        cam = cls(position=pos)
        cam.look_at(numpy.zeros(3))
        return cam

    @property
    def distance_from_origin(self) -> float:
        return float(numpy.linalg.norm(self.position))

    def basis(self):
        """(forward, right, up) (fly_camera.py:69-84)."""
        cp = math.cos(self.pitch)
        forward = numpy.array([cp * math.cos(self.yaw), cp * math.sin(self.yaw),
                               math.sin(self.pitch)])
        right = numpy.cross(forward, _WORLD_UP)
        norm = numpy.linalg.norm(right)
        if norm < 1e-9:
            right = numpy.array([math.sin(self.yaw), -math.cos(self.yaw), 0.0])
        else:
            right = right / norm
        return forward, right, numpy.cross(right, forward)

    def look_at(self, target) -> None:
        """Yaw and pitch toward ``target``, pitch clamped to +-89 deg
        (fly_camera.py:86-96)."""
        offset = numpy.asarray(target, dtype=float) - self.position
        norm = numpy.linalg.norm(offset)
        if norm < 1e-12:
            return
        d = offset / norm
        self.yaw = math.atan2(d[1], d[0])
        self.pitch = max(-_PITCH_LIMIT, min(_PITCH_LIMIT, math.asin(float(d[2]))))


@dataclass(frozen=True)
class OfflineSettings:
    """The renderer's physics and quality parameters (render.py:53-94)."""

    spin: float = 0.9
    fov_degrees: float = 70.0
    supersample: int = 2
    rtol: float = 1e-9
    refine_rtol: float = 1e-11
    max_steps: int = 12000
    max_step: float = 2.0
    capture_margin: float = 1e-3
    disk_enabled: bool = True
    disk_outer_radius: float = 18.0
    disk_temperature: float = 6500.0
    disk_detail: float = 1.0
    tile_rays: int = 120000
    backend: str = "hip"


def pinhole_grid(width: int, height: int, fov_degrees: float, supersample: int) -> numpy.ndarray:
    """Local (forward, right, up) sample directions on a supersample^2 grid
    per pixel, ordered by pixel row, column, subpixel row, column
    (render.py:101-144).  tan(fov/2) scales the horizontal coordinate."""
    tan_half = math.tan(math.radians(fov_degrees) / 2.0)
    aspect = width / height
    ss = supersample
    offsets = (numpy.arange(ss) + 0.5) / ss
    px = (numpy.arange(width)[:, None] + offsets[None, :]).reshape(-1)
    py = (numpy.arange(height)[:, None] + offsets[None, :]).reshape(-1)
    u = px / width * 2.0 - 1.0
    v = 1.0 - py / height * 2.0
    u_grid = numpy.tile(u.reshape(width, ss)[None, :, None, :], (height, 1, ss, 1))
    v_grid = numpy.tile(v.reshape(height, ss)[:, None, :, None], (1, width, 1, ss))
    local = numpy.stack([numpy.ones(u_grid.size), u_grid.reshape(-1) * tan_half,
                         v_grid.reshape(-1) * (tan_half / aspect)], axis=-1)
    return local / numpy.linalg.norm(local, axis=-1, keepdims=True)


def subpixel_directions(camera: FlyCamera, width: int, height: int, fov_degrees: float,
                        supersample: int) -> numpy.ndarray:
    """World-frame directions of every subpixel, (n, 3) (render.py:147-161)."""
    forward, right, up = camera.basis()
    local = pinhole_grid(width, height, fov_degrees, supersample)
    return (local[:, 0:1] * forward[None, :] + local[:, 1:2] * right[None, :]
            + local[:, 2:3] * up[None, :])


def pixel_tile_order(width: int, height: int, supersample: int, tile: int = 4) -> numpy.ndarray:
    """Permutation of subpixel rays listing tile x tile pixel blocks one after
    another (each pixel's supersample^2 rays together), so a 64-lane
    wavefront traces neighbouring pixels (4 x 4 pixels at supersample 2)."""
    ss2 = supersample * supersample
    r, c = numpy.meshgrid(numpy.arange(height), numpy.arange(width), indexing="ij")
    tiles_per_row = (width + tile - 1) // tile
    key = ((r // tile) * tiles_per_row + (c // tile)) * (tile * tile) + (r % tile) * tile + (c % tile)
    pix = numpy.argsort(key.ravel(), kind="stable")
    return (pix[:, None] * ss2 + numpy.arange(ss2)[None, :]).reshape(-1).astype(numpy.int64)


def disk_inner_radius(spacetime: KerrSpacetime) -> float:
    """ISCO of the +phi circular orbits the disk is built from
    (emission/novikov_thorne.py:31-37)."""
    return spacetime.isco_radius(prograde=spacetime.spin >= 0.0)


@dataclass
class TileResult:
    """Per-ray outputs of the offline trace (render.py:97-104) plus the
    accepted steps and attempts of the pass that produced the status."""

    status: Array
    escape_directions: Array
    hit_positions: Array
    hit_momenta: Array
    saturated: Array
    steps: Array
    attempts: Array
    iterations: int


def _launch_tile(spacetime, s0_soa, n, escape_radius, disk_radii, rtol, max_steps, max_step,
                 capture_margin, interior_stop, mode, tune, order, dev, states_out=None):
    if interior_stop is None:
        capture_radius = spacetime.outer_horizon_radius * (1.0 + capture_margin)
    else:
        capture_radius = float(interior_stop)
    m = resolve_mode(mode, rtol)
    p = _lib.TraceParams(spacetime.mass, spacetime.spin, float(escape_radius), capture_radius,
                         float(rtol), float(rtol) * 1e-3, 0.1, float(max_step), 0.0,
                         int(max_steps), _mode(m), 0, 0.0)
    flags = (_lib.TILE_DISK if disk_radii is not None else 0) | (
        _lib.TILE_INTERIOR if interior_stop is not None else 0)
    inner, outer = disk_radii if disk_radii is not None else (0.0, 0.0)
    t = _lib.TileParams(float(inner), float(outer), int(0.85 * max_steps), flags, 0)
    out = dict(status=torch.empty(n, dtype=torch.int32, device=dev),
               steps=torch.empty(n, dtype=torch.int64, device=dev),
               attempts=torch.empty(n, dtype=torch.int64, device=dev),
               saturated=torch.empty(n, dtype=torch.uint8, device=dev),
               escape_directions=torch.empty((n, 3), dtype=torch.float64, device=dev),
               hit_positions=torch.empty((n, 3), dtype=torch.float64, device=dev),
               hit_momenta=torch.empty((n, 3), dtype=torch.float64, device=dev),
               iters=torch.zeros(1, dtype=torch.int64, device=dev))
    lib = _lib.require_device()
    ld = s0_soa.shape[1]
    _lib.check(lib.bh_trace_tile_f64(
        p, t, (tune or Tune()).as_c(), _lib.ptr(s0_soa), ld, _lib.ptr(states_out), ld,
        _lib.ptr(out["status"]), _lib.ptr(out["steps"]), _lib.ptr(out["attempts"]),
        _lib.ptr(out["saturated"]), _lib.ptr(out["escape_directions"]),
        _lib.ptr(out["hit_positions"]), _lib.ptr(out["hit_momenta"]), _lib.ptr(out["iters"]),
        _lib.ptr(order), n, _lib.stream_handle()), "trace_tile")
    return out


def _to_soa(s0):
    lib = _lib.require_device()
    n = s0.shape[0]
    soa = torch.empty((8, max(n, 1)), dtype=torch.float64, device=s0.device)
    _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(s0), _lib.ptr(soa), soa.shape[1], n,
                                     _lib.stream_handle()), "aos_to_soa")
    return soa


def _result(out, ref) -> TileResult:
    conv = lambda t: _dev.like_input(t, ref)  # noqa: E731
    return TileResult(status=conv(out["status"]), escape_directions=conv(out["escape_directions"]),
                      hit_positions=conv(out["hit_positions"]),
                      hit_momenta=conv(out["hit_momenta"]),
                      saturated=conv(out["saturated"].bool()), steps=conv(out["steps"]),
                      attempts=conv(out["attempts"]), iterations=int(out["iters"].item()))


def trace_tile(spacetime: KerrSpacetime, state0: Array, escape_radius: float,
               settings: OfflineSettings, disk_radii: tuple[float, float] | None, rtol: float,
               max_steps: int, interior_stop: float | None = None, *, mode: str = "auto",
               tune: Tune | None = None, ray_order: Array | None = None) -> TileResult:
    """One pass of the offline tracing loop over (n, 8) states
    (render.py:184-309): status, flat escape directions, disk hits and the
    saturated flag, in the caller's array type.  Uses ``settings.max_step``
    and ``settings.capture_margin``; ``disk_radii`` = (inner, outer) or None;
    ``interior_stop`` switches to the interior-camera classification."""
    s0 = _dev.as_device_f64(state0, (8,), "state0")
    n = s0.shape[0]
    out = _launch_tile(spacetime, _to_soa(s0), n, escape_radius, disk_radii, rtol, max_steps,
                       settings.max_step, settings.capture_margin, interior_stop, mode, tune,
                       _order_tensor(ray_order, n, s0.device), s0.device)
    return _result(out, state0)


def trace_tile_retrace(spacetime: KerrSpacetime, state0: Array, escape_radius: float,
                       settings: OfflineSettings, disk_radii: tuple[float, float] | None,
                       interior_stop: float | None = None, *, mode: str = "auto",
                       tune: Tune | None = None, ray_order: Array | None = None):
    """The renderer's two passes (render.py:592-630): ``settings.rtol`` with
    ``settings.max_steps``, then the rays that saturated without hitting the
    disk or terminating, again at ``settings.refine_rtol`` with four times
    the budget; their status, escape direction and hit replace the first
    pass's (saturated, steps and attempts stay the first pass's).  The
    retraced rays are gathered on the device.  Returns (TileResult, retrace
    mask)."""
    s0 = _dev.as_device_f64(state0, (8,), "state0")
    n = s0.shape[0]
    dev = s0.device
    kw = dict(max_step=settings.max_step, capture_margin=settings.capture_margin,
              interior_stop=interior_stop, mode=mode, tune=tune)
    soa = _to_soa(s0)
    first = _launch_tile(spacetime, soa, n, escape_radius, disk_radii, settings.rtol,
                         settings.max_steps, order=_order_tensor(ray_order, n, dev), dev=dev,
                         **kw)
    retrace = first["saturated"].bool() & (first["status"] != int(RayStatus.DISK)) & (
        first["status"] != int(RayStatus.TERMINATED))
    idx = torch.nonzero(retrace).reshape(-1)
    if idx.numel():
        second = _launch_tile(spacetime, _to_soa(s0.index_select(0, idx).contiguous()),
                              idx.numel(), escape_radius, disk_radii, settings.refine_rtol,
                              settings.max_steps * 4, order=None, dev=dev, **kw)
        for name in ("status", "escape_directions", "hit_positions", "hit_momenta"):
            first[name].index_copy_(0, idx, second[name])
        first["iters"] = torch.maximum(first["iters"], second["iters"])
    return _result(first, state0), _dev.like_input(retrace, state0)


def camera_rays(spacetime: KerrSpacetime, camera: FlyCamera, width: int, height: int,
                settings: OfflineSettings) -> torch.Tensor:
    """(n, 8) past-directed photon states of every subpixel, built on the
    device (render.py:551-591 for a world camera)."""
    dirs = torch.from_numpy(subpixel_directions(camera, width, height, settings.fov_degrees,
                                                settings.supersample))
    dev = _dev.device()
    dirs = dirs.to(dev)
    pos = torch.from_numpy(numpy.asarray(camera.position, dtype=numpy.float64)).to(dev)
    pos = pos[None, :].expand(dirs.shape[0], 3).contiguous()
    mom = null_momentum_from_velocity(spacetime, pos, dirs, "past")
    return build_state(pos, mom)


def trace_frame(camera: FlyCamera, width: int, height: int, settings: OfflineSettings, *,
                mode: str = "auto", tune: Tune | None = None, tile: int | None = 4):
    """Trace a whole world-camera frame as the offline renderer does before
    its emission stage: disk annulus [ISCO of +phi orbits, max(outer,
    ISCO + 1)], escape radius 1.3 max(camera distance, disk outer), both
    passes; rays handed to wavefronts in ``tile`` x ``tile`` pixel blocks
    (None: the renderer's order).  Returns (TileResult with device tensors,
    retrace mask, initial states)."""
    st = KerrSpacetime(mass=1.0, spin=settings.spin)
    inner = disk_inner_radius(st)
    outer = max(settings.disk_outer_radius, inner + 1.0)
    disk = (inner, outer) if settings.disk_enabled else None
    escape_radius = 1.3 * max(camera.distance_from_origin, outer)
    s0 = camera_rays(st, camera, width, height, settings)
    order = None if tile is None else torch.from_numpy(
        pixel_tile_order(width, height, settings.supersample, tile))
    res, retrace = trace_tile_retrace(st, s0, escape_radius, settings, disk, mode=mode, tune=tune,
                                      ray_order=order)
    return res, retrace, s0
