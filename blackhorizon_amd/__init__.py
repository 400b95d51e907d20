"""blackhorizon_amd: BlackHorizon's batched adaptive Kerr geodesic tracer,
re-built for AMD Instinct MI355X (gfx950) with hand-written FP64 HIP kernels.

Public surface mirrors the reference package (blackhorizon/__init__.py:9-24).
"""

from .backend import gpu_available, get_xp, to_numpy, xp_of
from .camera import PinholeCamera
from .kerr import KerrSpacetime, KerrSchildGeometry
from .tracer import RayStatus, TraceResult, Tune, trace_rays

__version__ = "0.1.0"

__all__ = [
    "KerrSpacetime",
    "KerrSchildGeometry",
    "PinholeCamera",
    "RayStatus",
    "TraceResult",
    "Tune",
    "trace_rays",
    "gpu_available",
    "get_xp",
    "xp_of",
    "to_numpy",
    "__version__",
]
