"""The C-ABI library loads, exports every symbol include/bh_trace.h
declares, and rejects invalid arguments before touching a device."""

import ctypes

import pytest

from blackhorizon_amd import _lib


def test_library_exports_header_symbols():
    lib = _lib.load()
    declared = _lib.header_symbols()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.SIGNATURES)
    assert lib.bh_abi_version() == _lib.ABI_VERSION


def _params(**over):
    p = dict(mass=1.0, spin=0.9, escape_radius=45.0, capture_radius=1.5, rtol=1e-8, atol=1e-10,
             initial_step=0.1, max_step=2.0, min_step=1e-10, max_steps=20000, mode=0, reserved=0,
             lambda_end=float("inf"))
    p.update(over)
    return _lib.TraceParams(**p)


@pytest.mark.parametrize("over,msg", [
    (dict(spin=1.01), "Kerr bound"),
    (dict(mass=-1.0), "mass must be positive"),
    (dict(mode=7), "mode"),
])
def test_invalid_params_rejected(over, msg):
    lib = _lib.load()
    rc = lib.bh_trace_plain_f64(_params(**over), None, None, 0, None, 0, None, None, None, None,
                                None, 0, None)
    assert rc == _lib.BH_EINVAL
    assert msg in lib.bh_last_error().decode()
    with pytest.raises(ValueError):
        _lib.check(rc, "trace")


@pytest.mark.parametrize("max_steps", [(1 << 60) + 1, -(1 << 60) - 1, 1 << 62, -(1 << 63)])
def test_max_steps_range_rejected(max_steps):
    """|max_steps| beyond 2^60 would overflow the attempt cap 4 * max_steps:
    the ABI refuses it for one job, and the host refuses it for every job
    (the batched launch reads its job table on the device)."""
    from blackhorizon_amd import batch, tracer
    from blackhorizon_amd.kerr import KerrSpacetime

    lib = _lib.load()
    rc = lib.bh_trace_plain_f64(_params(max_steps=max_steps), None, None, 0, None, 0, None, None,
                                None, None, None, 0, None)
    assert rc == _lib.BH_EINVAL
    assert "max_steps out of range" in lib.bh_last_error().decode()
    st = KerrSpacetime(spin=0.5)
    with pytest.raises(ValueError, match="max_steps"):
        tracer.make_params(st, 45.0, max_steps, 1e-8, 1e-10, 0.1, 2.0, 1e-10, 1e-3)
    with pytest.raises(ValueError, match="max_steps"):
        batch._Group([batch.TraceJob(st, None, 45.0, max_steps=max_steps)], [0], "fast", None,
                     None)


def test_negative_max_steps_accepted():
    """A negative max_steps is valid (the reference's loop then runs no
    iteration): only the range is refused."""
    lib = _lib.load()
    rc = lib.bh_trace_plain_f64(_params(max_steps=-5), None, None, 0, None, 0, None, None, None,
                                None, None, 0, None)
    assert rc == _lib.BH_OK


def test_invalid_tune_rejected():
    lib = _lib.load()
    for tune in (_lib.TuneParams(96, 1, 1, 0), _lib.TuneParams(512, 1, 1, 0),
                 _lib.TuneParams(256, -1, 1, 0), _lib.TuneParams(256, 1, 65, 0),
                 _lib.TuneParams(256, 1, 1, 4), _lib.TuneParams(256, 1, 1, -1)):
        rc = lib.bh_trace_plain_f64(_params(), tune, None, 0, None, 0, None, None, None, None,
                                    None, 0, None)
        assert rc == _lib.BH_EINVAL


@pytest.mark.parametrize("mode,tune,msg", [
    (_lib.MODE_FAST, (256, 1, 1, 4), "needs the precise build"),
    (_lib.MODE_PRECISE, (256, 1, 1, 2), "lanes_per_ray must be 1 or 4"),
    (_lib.MODE_PRECISE, (256, 1, 1, -4), "lanes_per_ray must be 1 or 4"),
    (_lib.MODE_PRECISE, (256, 1, 17, 4), "refill_min must be in [1, 16]"),
])
def test_batched_quad_tune_rejected(mode, tune, msg):
    """The quad tracer (lanes_per_ray 4) exists in the precise build only
    and takes at most 16 idle quads per refill; other lane counts are
    refused before any device work."""
    lib = _lib.load()
    jobs = (_lib.TraceParams * 1)(_params())
    dummy = ctypes.c_void_p(8)  # never dereferenced: the call fails its checks first
    rc = lib.bh_trace_batched_f64(jobs, dummy, 1, mode, _lib.TuneParams(*tune), None, 0, None, 0,
                                  None, None, None, dummy, None, dummy, 0, None)
    assert rc == _lib.BH_EINVAL
    assert msg in lib.bh_last_error().decode()


def test_bad_sizes_rejected():
    lib = _lib.load()
    assert lib.bh_trace_plain_f64(_params(), None, None, 5, None, 5, None, None, None, None,
                                  None, 10, None) == _lib.BH_EINVAL
    assert lib.bh_trace_batched_f64(None, None, 0, 0, None, None, 0, None, 0, None, None, None,
                                    None, None, None, 0, None) == _lib.BH_EINVAL
    assert lib.bh_trace_batched_f64(None, None, 300, 0, None, None, 0, None, 0, None, None, None,
                                    None, None, None, 0, None) == _lib.BH_EINVAL
    assert lib.bh_aos_to_soa_f64(None, None, 3, 4, None) == _lib.BH_EINVAL
    assert lib.bh_geometry_f64(1.0, 2.0, 0, None, 0, None, None, None, None, None, None) \
        == _lib.BH_EINVAL


def test_python_api_errors_without_device(monkeypatch):
    import torch

    import blackhorizon_amd as bh

    if torch.cuda.is_available():
        pytest.skip("device present")
    with pytest.raises(RuntimeError):
        bh.trace_rays(bh.KerrSpacetime(spin=0.5), [[0.0, 10.0, 0, 0, 1.0, -1.0, 0, 0]], 20.0)
    with pytest.raises(ValueError):
        bh.KerrSpacetime(spin=1.5)
    assert not bh.gpu_available()
    with pytest.raises(RuntimeError):
        bh.get_xp("hip")
    with pytest.raises(ValueError):
        bh.get_xp("cuda")


def test_new_entries_reject_bad_arguments():
    """The offline, shader, dense, fixed-step, camera, shading and resource
    entries validate before any device work."""
    lib = _lib.load()
    E = _lib.BH_EINVAL
    tile = _lib.TileParams(6.0, 2.0, 10, _lib.TILE_DISK, 0)  # inner > outer
    assert lib.bh_trace_tile_f64(_params(), tile, None, None, 0, None, 0, None, None, None, None,
                                 None, None, None, None, None, 0, None) == E
    assert lib.bh_trace_tile_f64(_params(), _lib.TileParams(2.0, 6.0, 10, 8, 0), None, None, 0,
                                 None, 0, None, None, None, None, None, None, None, None, None, 0,
                                 None) == E  # unknown flag
    assert lib.bh_trace_tile_f64(_params(), None, None, None, 0, None, 0, None, None, None, None,
                                 None, None, None, None, None, 0, None) == E
    sp = _lib.ShaderParams(1.0, 0.9, 1.5, 60.0, 0.22, 1e-4, 4.0, 1e3, 0.0, 0.0, 512, 0, 9)
    assert lib.bh_trace_shader_f64(sp, None, 0, None, 0, None, None, None, 0, None) == E
    assert lib.bh_integrate_dense_f64(1.0, 0.9, 1, 1e-10, 1e-12, 2.0, 1.0, None, 0, 0, None, None,
                                      1, None, None, None, None) == E  # n = 0
    assert lib.bh_rk4_f64(1.0, 0.9, 1, None, 0, None, 4, 1, None) == E  # ld < n
    assert lib.bh_midpoint_f64(1.0, 0.9, 1, None, 4, None, 4, 1, -1, None) == E
    assert lib.bh_camera_rays_f64(1.0, 0.9, None, 0.1, 0.1, 4, 4, 1, None, 16, None) == E
    assert lib.bh_shade_checker_u8(1.0, 0.9, None, 4, None, 4, 0, None, None) == E  # cells 0
    out = (ctypes.c_int32 * 4)()
    assert lib.bh_kernel_attributes(9, 0, 256, ctypes.addressof(out)) == E
    assert lib.bh_kernel_attributes(0, 0, 96, ctypes.addressof(out)) == E
    assert lib.bh_adjoint_scratch_doubles(0) == -1
    assert lib.bh_adjoint_scratch_doubles(32) == 9 * 32


def test_controller_and_colour_entries_reject_bad_arguments():
    """error_ratio / step_factor / checker colours refuse unknown modes, bad
    counts, missing buffers and a zero cell count before any device work."""
    lib = _lib.load()
    E = _lib.BH_EINVAL
    buf = (ctypes.c_double * 8)()
    p = ctypes.addressof(buf)
    assert lib.bh_error_ratio_f64(7, p, p, p, 1, 8, 1e-8, 1e-10, p, None) == E
    assert "mode" in lib.bh_last_error().decode()
    assert lib.bh_error_ratio_f64(0, p, p, p, -1, 8, 1e-8, 1e-10, p, None) == E
    assert lib.bh_error_ratio_f64(0, p, p, p, 1, 0, 1e-8, 1e-10, p, None) == E
    assert lib.bh_error_ratio_f64(1, p, p, p, 1, 129, 1e-8, 1e-10, p, None) == E
    assert "dim" in lib.bh_last_error().decode()
    assert lib.bh_error_ratio_f64(0, p, None, p, 1, 8, 1e-8, 1e-10, p, None) == E
    assert lib.bh_step_factor_f64(3, p, 1, p, None) == E
    assert lib.bh_step_factor_f64(0, p, 1 << 40, p, None) == E
    assert lib.bh_step_factor_f64(1, None, 1, p, None) == E
    rgb = (ctypes.c_uint8 * 3)()
    assert lib.bh_checker_colors_u8(p, 1, 0, ctypes.addressof(rgb), None) == E
    assert lib.bh_checker_colors_u8(p, -2, 12, ctypes.addressof(rgb), None) == E
    assert lib.bh_checker_colors_u8(p, 1, 12, None, None) == E
    # n = 0 is a no-op that succeeds without a device
    assert lib.bh_error_ratio_f64(0, None, None, None, 0, 2, 1e-8, 1e-10, None, None) == 0
    assert lib.bh_step_factor_f64(1, None, 0, None, None) == 0
    assert lib.bh_checker_colors_u8(None, 0, 12, None, None) == 0


def test_static_scratch_adjoint_refuses_beyond_its_limit():
    """The allocation-free reverse pass refuses, before touching a device,
    calls whose scratch (9 K doubles per ray, rays rounded up to 64) exceeds
    its static array, and the arguments the allocating entry refuses."""
    lib = _lib.load()
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    limit = 8 * 1024 * 1024
    args = lambda n, K: (_params(), n, K, 4.0, p, p, p, p, p, p, p, p, p, None, None)  # noqa
    # 29,120 rays at K = 32 fit (9 * 32 * 29120 <= 2^23); 29,184 (the next multiple of 64) do not
    assert 9 * 32 * 29120 <= limit < 9 * 32 * 29184
    rc = lib.bh_trace_adjoint_static_f64(*args(29121, 32))
    assert rc == _lib.BH_EINVAL
    assert "static scratch" in lib.bh_last_error().decode()
    rc = lib.bh_trace_adjoint_static_f64(*args(1 << 20, 1))
    assert rc == _lib.BH_EINVAL and "static scratch" in lib.bh_last_error().decode()
    assert lib.bh_trace_adjoint_static_f64(*args(64, 0)) == _lib.BH_EINVAL
    assert "K <= 2^20" in lib.bh_last_error().decode()
    rc = lib.bh_trace_adjoint_static_f64(_params(), 64, 8, 4.0, None, p, p, p, p, p, p, p, p,
                                         None, None)
    assert rc == _lib.BH_EINVAL and "NULL" in lib.bh_last_error().decode()
    # nothing to reverse: no device work at all
    assert lib.bh_trace_adjoint_static_f64(*args(0, 8)) == _lib.BH_OK
