"""Pinhole camera (reference camera.py:24-120).

Host-side, like the reference: it maps pixels to world directions with
NumPy; photon momenta are then built on the device by
``geodesics.null_momentum_from_velocity``.  Horizontal FOV; pixel (row, col)
is ray ``row * width + col`` from the top-left corner.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy


def _normalize(v: numpy.ndarray) -> numpy.ndarray:
    norm = numpy.linalg.norm(v)
    if norm == 0.0:
        raise ValueError("zero-length vector cannot be normalized")
    return v / norm


@dataclass(frozen=True)
class PinholeCamera:
    position: tuple
    target: tuple = (0.0, 0.0, 0.0)
    up_hint: tuple = (0.0, 0.0, 1.0)
    fov_degrees: float = 20.0
    width: int = 320
    height: int = 240

    forward: numpy.ndarray = field(init=False, repr=False)
    right: numpy.ndarray = field(init=False, repr=False)
    up: numpy.ndarray = field(init=False, repr=False)

    def __post_init__(self) -> None:
        # camera.py:48-66
        if self.width <= 0 or self.height <= 0:
            raise ValueError("image dimensions must be positive")
        if not 0.0 < self.fov_degrees < 180.0:
            raise ValueError("fov_degrees must be in (0, 180)")
        forward = _normalize(numpy.asarray(self.target, dtype=numpy.float64)
                             - numpy.asarray(self.position, dtype=numpy.float64))
        right = numpy.cross(forward, numpy.asarray(self.up_hint, dtype=numpy.float64))
        if numpy.linalg.norm(right) < 1e-12:
            right = numpy.cross(forward, numpy.array([0.0, 1.0, 0.0]))
        right = _normalize(right)
        up = numpy.cross(right, forward)
        object.__setattr__(self, "forward", forward)
        object.__setattr__(self, "right", right)
        object.__setattr__(self, "up", up)

    @classmethod
    def from_orbit(cls, distance: float, inclination_degrees: float,
                   azimuth_degrees: float = 0.0, **kwargs) -> "PinholeCamera":
        """Camera on a sphere of radius ``distance`` (camera.py:68-94)."""
        if distance <= 0.0:
            raise ValueError("distance must be positive")
        inc = math.radians(inclination_degrees)
        azi = math.radians(azimuth_degrees)
        position = (distance * math.sin(inc) * math.cos(azi),
                    distance * math.sin(inc) * math.sin(azi),
                    distance * math.cos(inc))
        return cls(position=position, **kwargs)

    def ray_directions(self) -> numpy.ndarray:
        """Unit directions, (height * width, 3), row-major (camera.py:96-113)."""
        half_w = math.tan(math.radians(self.fov_degrees) / 2.0)
        half_h = half_w * self.height / self.width
        cols = (numpy.arange(self.width) + 0.5) / self.width * 2.0 - 1.0
        rows = 1.0 - (numpy.arange(self.height) + 0.5) / self.height * 2.0
        u, v = numpy.meshgrid(cols * half_w, rows * half_h)
        d = (self.forward[None, None, :] + u[..., None] * self.right[None, None, :]
             + v[..., None] * self.up[None, None, :])
        d /= numpy.linalg.norm(d, axis=-1, keepdims=True)
        return d.reshape(-1, 3)

    def ray_origins(self) -> numpy.ndarray:
        """Camera position per pixel, (height * width, 3) (camera.py:115-120)."""
        o = numpy.asarray(self.position, dtype=numpy.float64)
        return numpy.broadcast_to(o, (self.height * self.width, 3)).copy()

    def device_states(self, spacetime, time_orientation: str = "past"):
        """(height * width, 8) photon states of every pixel built on the
        device in one kernel -- the directions of ray_directions(), unit
        |p_t| covectors (null_momentum_from_velocity) and build_state at
        t = 0 -- without host arrays or a host-to-device copy.  Equal to
        those three steps done separately (bitwise)."""
        import torch

        from . import _dev, _lib

        if time_orientation not in ("past", "future"):
            raise ValueError("time_orientation must be 'past' or 'future'")
        lib = _lib.require_device()
        n = self.width * self.height
        dev = _dev.device()
        soa = torch.empty((8, n), dtype=torch.float64, device=dev)
        half_w = math.tan(math.radians(self.fov_degrees) / 2.0)
        half_h = half_w * self.height / self.width
        cam = numpy.concatenate([numpy.asarray(self.position, dtype=numpy.float64),
                                 self.forward, self.right, self.up]).astype(numpy.float64)
        stream = _lib.stream_handle()
        _lib.check(lib.bh_camera_rays_f64(spacetime.mass, spacetime.spin,
                                          cam.ctypes.data, half_w, half_h, self.width,
                                          self.height, 1 if time_orientation == "past" else 0,
                                          _lib.ptr(soa), n, stream), "camera_rays")
        out = torch.empty((n, 8), dtype=torch.float64, device=dev)
        _lib.check(lib.bh_soa_to_aos_f64(_lib.ptr(soa), n, _lib.ptr(out), n, stream),
                   "soa_to_aos")
        return out

    def tile_order(self, tile: int = 8) -> numpy.ndarray:
        """Permutation of ray indices that lists the pixels tile by tile
        (tile x tile blocks, row-major inside and across tiles; partial
        tiles at the right and bottom edges included).  Handing rays to
        64-lane wavefronts in this order groups neighbouring pixels, whose
        trace lengths are similar, into the same wavefront."""
        if tile < 1:
            raise ValueError("tile must be >= 1")
        rows = numpy.arange(self.height)
        cols = numpy.arange(self.width)
        r, c = numpy.meshgrid(rows, cols, indexing="ij")
        tiles_per_row = (self.width + tile - 1) // tile
        key = ((r // tile) * tiles_per_row + (c // tile)) * (tile * tile) \
            + (r % tile) * tile + (c % tile)
        return numpy.argsort(key.ravel(), kind="stable").astype(numpy.int64)

    def center_tile_order(self, tile: int = 8) -> numpy.ndarray:
        """tile_order with the tiles sorted by the distance of their centre
        from the image centre, nearest first (ties in row-major order).  A
        camera aimed at the hole sees the shadow and the photon ring -- the
        longest traces -- around the image centre, so their wavefronts are
        dispatched first and the kernel's tail is made of short rays (the
        longest-job-first rule, with no trace needed to rank the tiles)."""
        if tile < 1:
            raise ValueError("tile must be >= 1")
        tiles_per_row = (self.width + tile - 1) // tile
        tiles_per_col = (self.height + tile - 1) // tile
        ty, tx = numpy.meshgrid(numpy.arange(tiles_per_col), numpy.arange(tiles_per_row),
                                indexing="ij")
        d = numpy.hypot((ty + 0.5) * tile - self.height / 2.0,
                        (tx + 0.5) * tile - self.width / 2.0).ravel()
        rank = numpy.argsort(d, kind="stable")
        rank_of = numpy.empty_like(rank)
        rank_of[rank] = numpy.arange(rank.size)
        r, c = numpy.meshgrid(numpy.arange(self.height), numpy.arange(self.width), indexing="ij")
        tid = (r // tile) * tiles_per_row + (c // tile)
        key = rank_of[tid] * (tile * tile) + (r % tile) * tile + (c % tile)
        return numpy.argsort(key.ravel(), kind="stable").astype(numpy.int64)

