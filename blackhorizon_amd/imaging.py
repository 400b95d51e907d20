"""Stage-1 shading of traced rays (reference imaging.py).

``render_image`` turns a TraceResult into the reference's RGB image on the
device (``bh_shade_checker_u8``): captured rays black, escaped rays on a
12-cell celestial checkerboard of their final coordinate-velocity
direction, MAX_STEPS dark red, anything else magenta (imaging.py:20-91).
"""

from __future__ import annotations

import numpy
import torch

from . import _dev, _lib
from .backend import to_numpy
from .kerr import KerrSpacetime
from .tracer import TraceResult

COLOR_SHADOW = (0, 0, 0)
COLOR_CHECKER_DARK = (24, 30, 52)
COLOR_CHECKER_LIGHT = (104, 128, 178)
COLOR_MAX_STEPS = (120, 32, 32)
COLOR_FAILED = (200, 0, 200)


def celestial_checkerboard(directions, cells: int = 12) -> numpy.ndarray:
    """RGB uint8 (n, 3) checker colours of unit directions (n, 3) on the
    celestial sphere, ``cells`` cells over the polar angle (imaging.py:27-50)."""
    d = _dev.as_device_f64(directions, (3,), "directions")
    n = d.shape[0]
    rgb = torch.empty((n, 3), dtype=torch.uint8, device=d.device)
    lib = _lib.require_device()
    _lib.check(lib.bh_checker_colors_u8(_lib.ptr(d), n, int(cells), _lib.ptr(rgb),
                                        _lib.stream_handle()), "celestial_checkerboard")
    return to_numpy(rgb)


def render_image(spacetime: KerrSpacetime, result: TraceResult, width: int, height: int,
                 cells: int = 12) -> numpy.ndarray:
    """RGB uint8 image (height, width, 3) of ``height * width`` traced rays
    in row-major pixel order (imaging.py:53-91)."""
    n = width * height
    states = _dev.as_device_f64(result.states, (8,), "states")
    if states.shape[0] != n:
        raise ValueError("result size does not match image dimensions")
    status = torch.as_tensor(result.status).to(states.device, torch.int32).contiguous()
    lib = _lib.require_device()
    stream = _lib.stream_handle()
    soa = torch.empty((8, max(n, 1)), dtype=torch.float64, device=states.device)
    _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(states), _lib.ptr(soa), soa.shape[1], n, stream),
               "aos_to_soa")
    rgb = torch.empty((n, 3), dtype=torch.uint8, device=states.device)
    _lib.check(lib.bh_shade_checker_u8(spacetime.mass, spacetime.spin, _lib.ptr(soa),
                                       soa.shape[1], _lib.ptr(status), n, int(cells),
                                       _lib.ptr(rgb), stream), "shade_checker")
    return to_numpy(rgb).reshape(height, width, 3)


def save_png(image: numpy.ndarray, path: str) -> None:
    """Write an RGB uint8 (h, w, 3) array to a PNG file (imaging.py:94-96)."""
    from PIL import Image

    Image.fromarray(numpy.ascontiguousarray(image), mode="RGB").save(path)
