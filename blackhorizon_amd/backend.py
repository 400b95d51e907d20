"""Array-backend dispatch with a third, HIP backend (reference backend.py).

The reference maps backend names to array modules: ``"cpu"`` -> NumPy and
``"gpu"`` -> CuPy (backend.py:34-45), and asks ``xp_of(array)`` which module
owns an array (48-52).  Here the device arrays are PyTorch-ROCm tensors and
the compute behind them is this package's HIP kernels, so:

* ``get_xp("hip")`` returns ``torch`` (RuntimeError without a HIP device);
  ``get_xp("cpu")`` returns ``numpy`` for host arrays; any other name raises
  ValueError, as backend.py:45 does.
* ``xp_of`` recognises torch tensors.
* ``to_numpy`` copies device tensors to host (backend.py:55-59).
* ``gpu_available`` reports a usable HIP device with the library loaded.
"""

from __future__ import annotations

from typing import Any

import numpy

from . import _lib

Array = Any
"""A NumPy ndarray or a PyTorch-ROCm tensor."""

BACKENDS = ("cpu", "hip")


def gpu_available() -> bool:
    """True if a HIP device is visible and libbhtrace loads (backend.py:24-31)."""
    return _lib.device_ready()


def get_xp(backend: str):
    """Array module for a backend name (backend.py:34-45)."""
    if backend == "cpu":
        return numpy
    if backend in ("hip", "gpu"):
        if not gpu_available():
            raise RuntimeError(
                "HIP backend requested but no ROCm device is visible or "
                "libbhtrace is not built (python -m blackhorizon_amd.build)")
        import torch

        return torch
    raise ValueError(f"Unknown backend {backend!r}, expected 'cpu' or 'hip'.")


def is_tensor(array: Array) -> bool:
    import torch

    return isinstance(array, torch.Tensor)


def xp_of(array: Array):
    """The module owning ``array`` (backend.py:48-52)."""
    if is_tensor(array):
        import torch

        return torch
    return numpy


def to_numpy(array: Array) -> numpy.ndarray:
    """Host copy as a NumPy array (backend.py:55-59); device tensors come
    back through pinned host memory."""
    if is_tensor(array):
        from . import _dev

        return _dev.to_numpy(array)
    return numpy.asarray(array)
