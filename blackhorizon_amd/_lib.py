"""ctypes binding of ``libbhtrace.so`` (C ABI in include/bh_trace.h).

The library is linked against ``libamdhip64.so.7``; PyTorch-ROCm is imported
first so that the dynamic loader binds the library to the HIP runtime torch
has already loaded (same soname), and device pointers and streams from torch
tensors are valid inside the library.
"""

from __future__ import annotations

import ctypes
import os
import re

import torch  # noqa: F401  (must precede loading the library: HIP runtime binding)

# BH_TRACE_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("BH_TRACE_LIB") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "lib", "libbhtrace.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "include", "bh_trace.h")

BH_OK = 0
BH_EINVAL = -1
BH_EHIP = -2
MODE_FAST = 0
MODE_PRECISE = 1
DENSE_DONE = 0
DENSE_CHUNK_FULL = 1
DENSE_UNDERFLOW = 2
ABI_VERSION = 4


class TraceParams(ctypes.Structure):
    """bh_trace_params."""

    _fields_ = [
        ("mass", ctypes.c_double),
        ("spin", ctypes.c_double),
        ("escape_radius", ctypes.c_double),
        ("capture_radius", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("initial_step", ctypes.c_double),
        ("max_step", ctypes.c_double),
        ("min_step", ctypes.c_double),
        ("max_steps", ctypes.c_int64),
        ("mode", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("lambda_end", ctypes.c_double),
    ]


class TileParams(ctypes.Structure):
    """bh_tile_params."""

    _fields_ = [
        ("disk_inner", ctypes.c_double),
        ("disk_outer", ctypes.c_double),
        ("saturate_steps", ctypes.c_int64),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ShaderParams(ctypes.Structure):
    """bh_shader_params."""

    _fields_ = [(name, ctypes.c_double) for name in (
        "mass", "spin", "capture_radius", "escape_radius", "step_scale", "min_step", "max_step",
        "momentum_bailout", "disk_inner", "disk_outer")] + [
        ("max_steps", ctypes.c_int64), ("flags", ctypes.c_int32), ("mode", ctypes.c_int32)]


TILE_DISK = 1
TILE_INTERIOR = 2


class TuneParams(ctypes.Structure):
    """bh_tune_params."""

    _fields_ = [
        ("block_size", ctypes.c_int32),
        ("rays_per_lane", ctypes.c_int32),
        ("refill_min", ctypes.c_int32),
        ("lanes_per_ray", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double

# name -> argtypes (restype is int32 unless listed in _RESTYPES)
SIGNATURES = {
    "bh_abi_version": [],
    "bh_last_error": [],
    "bh_count_outside_f64": [ctypes.POINTER(TraceParams), _P, _I64, _I64, _P, _P],
    "bh_trace_plain_f64": [ctypes.POINTER(TraceParams), ctypes.POINTER(TuneParams), _P, _I64,
                           _P, _I64, _P, _P, _P, _P, _P, _I64, _P],
    "bh_trace_batched_f64": [_P, _P, _I32, _I32, ctypes.POINTER(TuneParams), _P, _I64, _P, _I64,
                             _P, _P, _P, _P, _P, _P, _I64, _P],
    "bh_adjoint_scratch_doubles": [_I32],
    "bh_trace_record_f64": [ctypes.POINTER(TraceParams), _P, _I64, _P, _I64, _P, _P, _P, _P, _P,
                            _I64, _I32, _I32, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "bh_trace_record_trajectory_f64": [ctypes.POINTER(TraceParams), _P, _I64, _P, _I64, _P, _P,
                                       _P, _P, _P, _I64, _P, _I64, _P, _P, _D, _P, _P, _P, _P,
                                       _P, _P, _P, _P],
    "bh_trace_adjoint_trajectory_f64": [ctypes.POINTER(TraceParams), _I64, _D, _P, _I64, _P, _P,
                                        _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "bh_trace_adjoint_f64": [ctypes.POINTER(TraceParams), _I64, _I32, _D, _P, _P, _P, _P, _P, _P,
                             _P, _P, _P, _P, _P, _P],
    "bh_trace_adjoint_static_f64": [ctypes.POINTER(TraceParams), _I64, _I32, _D, _P, _P, _P, _P,
                                    _P, _P, _P, _P, _P, _P, _P],
    "bh_rhs_vjp_f64": [_D, _D, _P, _P, _I64, _P, _P, _P],
    "bh_radial_minimum_f64": [_D, _P, _P, _P, _P, _I64, _I32, _I32, _P, _P, _P],
    "bh_aos_to_soa_f64": [_P, _P, _I64, _I64, _P],
    "bh_soa_to_aos_f64": [_P, _I64, _P, _I64, _P],
    "bh_geometry_f64": [_D, _D, _I32, _P, _I64, _P, _P, _P, _P, _P, _P],
    "bh_geodesic_rhs_f64": [_D, _D, _I32, _P, _I64, _P, _P],
    "bh_dp5_step_f64": [_D, _D, _I32, _D, _D, _P, _P, _I64, _P, _P, _P, _P, _P],
    "bh_error_ratio_f64": [_I32, _P, _P, _P, _I64, _I32, _D, _D, _P, _P],
    "bh_step_factor_f64": [_I32, _P, _I64, _P, _P],
    "bh_checker_colors_u8": [_P, _I64, _I32, _P, _P],
    "bh_trace_tile_f64": [ctypes.POINTER(TraceParams), ctypes.POINTER(TileParams),
                          ctypes.POINTER(TuneParams), _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P,
                          _P, _P, _P, _I64, _P],
    "bh_camera_rays_f64": [_D, _D, _P, _D, _D, _I64, _I64, _I32, _P, _I64, _P],
    "bh_shade_checker_u8": [_D, _D, _P, _I64, _P, _I64, _I32, _P, _P],
    "bh_kernel_attributes": [_I32, _I32, _I32, _P],
    "bh_trace_shader_f64": [ctypes.POINTER(ShaderParams), _P, _I64, _P, _I64, _P, _P, _P, _I64,
                            _P],
    "bh_midpoint_f64": [_D, _D, _I32, _P, _I64, _P, _I64, _I64, _I32, _P],
    "bh_rk4_f64": [_D, _D, _I32, _P, _I64, _P, _I64, _I64, _P],
    "bh_integrate_dense_f64": [_D, _D, _I32, _D, _D, _D, _D, _P, _I64, _I64, _P, _P, _I64, _P,
                               _P, _P, _P],
    "bh_null_momentum_f64": [_D, _D, _P, _P, _I64, _I32, _I32, _P, _P],
    "bh_timelike_momentum_f64": [_D, _D, _P, _P, _I64, _P, _P, _P],
    "bh_diagnostics_f64": [_D, _D, _P, _I64, _P, _P, _P, _P, _P],
}
_RESTYPES = {"bh_last_error": ctypes.c_char_p}

_lib = None


def header_symbols(path: str = HEADER_PATH) -> list[str]:
    """Function names declared in the public header."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char \*)\s*(bh_\w+)\s*\(", text, re.M)))


def load() -> ctypes.CDLL:
    """Load and type the library (no device work is done)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m blackhorizon_amd.build` "
            "(hipcc, gfx950) first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, ctypes.c_int32)
    if lib.bh_abi_version() != ABI_VERSION:
        raise RuntimeError("libbhtrace ABI version mismatch; rebuild the library")
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    """Map a BH_* return code onto the reference's exception types."""
    if rc == BH_OK:
        return
    msg = load().bh_last_error().decode(errors="replace")
    if rc == BH_EINVAL:
        raise ValueError(msg or what)
    raise RuntimeError(f"{what}: {msg}")


def device_ready() -> bool:
    """True when a HIP device is visible to torch and the library loads
    (an unbuilt, unloadable or mismatched library counts as not ready)."""
    if not torch.cuda.is_available():
        return False
    try:
        load()
    except (OSError, RuntimeError, AttributeError):
        return False
    return True


def require_device() -> ctypes.CDLL:
    """The library, or a loud error when there is no usable HIP device."""
    lib = load()
    if not torch.cuda.is_available():
        raise RuntimeError("blackhorizon_amd needs a HIP (ROCm) GPU; none is visible")
    return lib


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


KERNELS = {"plain": 0, "batched": 1, "record": 2, "adjoint": 3, "tile": 4, "shader": 5,
           "quad": 6}


def kernel_attributes(kernel: str, mode: str = "fast", block: int = 256) -> dict:
    """Registers, static LDS and spill scratch per thread of one tracing
    kernel, and its resident waves per SIMD at ``block`` threads per block
    (bh_kernel_attributes)."""
    lib = require_device()
    out = (ctypes.c_int32 * 4)()
    check(lib.bh_kernel_attributes(KERNELS[kernel], MODE_FAST if mode == "fast" else MODE_PRECISE,
                                   block, ctypes.addressof(out)), "kernel_attributes")
    return dict(vgprs=out[0], lds_bytes=out[1], scratch_bytes=out[2], blocks_per_cu=out[3],
                waves_per_simd=out[3] * (block // 64) / 4.0)


def ptr(t) -> int | None:
    """Device address of a tensor (None for None)."""
    return None if t is None else int(t.data_ptr())
