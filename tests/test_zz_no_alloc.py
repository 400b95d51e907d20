"""GPU: the allocation-free reverse pass (bh_trace_adjoint_static_f64, its
scratch in a static device array) against the allocating one
(bh_trace_adjoint_f64, caller scratch), both called through the C ABI on the
same recorded forward: bit-identical gradients on two calls in a row, the
second with fewer rays, in both arithmetic builds and lane orders."""

import math

import numpy as np
import pytest
import torch

from oracle import bh_oracle as orc

pytestmark = pytest.mark.gpu

SPIN = 0.7


def _rays(width, height):
    st = orc.Kerr(spin=SPIN)
    cam = orc.PinholeCamera.from_orbit(40.0, 45.0, fov_degrees=30.0, width=width, height=height)
    return torch.from_numpy(orc.camera_state(st, cam)).cuda()


def _both(rec, gterm_rows, order):
    """(allocating, static) reverse passes of one recorded forward."""
    from blackhorizon_amd import _lib

    lib = _lib.require_device()
    n = rec.n
    dev = rec.term.device
    gterm = torch.zeros((8, n), dtype=torch.float64, device=dev)
    gterm[:, :n] = gterm_rows.t()
    stream = _lib.stream_handle()
    out = []
    for static in (False, True):
        gy0 = torch.full((8, n), float("nan"), dtype=torch.float64, device=dev)
        gspin = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
        common = (rec.params, n, rec.cfg.K, rec.score_radius, _lib.ptr(rec.ckpt),
                  _lib.ptr(rec.event), _lib.ptr(rec.ev_step), _lib.ptr(rec.evy1),
                  _lib.ptr(rec.prev), _lib.ptr(rec.theta), _lib.ptr(gterm), _lib.ptr(gy0),
                  _lib.ptr(gspin))
        if static:
            rc = lib.bh_trace_adjoint_static_f64(*common, _lib.ptr(order), stream)
        else:
            per_ray = lib.bh_adjoint_scratch_doubles(rec.cfg.K)
            scratch = torch.empty(per_ray * (-(-n // 64) * 64), dtype=torch.float64, device=dev)
            rc = lib.bh_trace_adjoint_f64(*common, _lib.ptr(scratch), _lib.ptr(order), stream)
        _lib.check(rc, "adjoint")
        torch.cuda.synchronize()
        out.append((gy0.cpu().numpy(), gspin.cpu().numpy()))
    return out


@pytest.mark.parametrize("mode", ["fast", "precise"])
def test_static_scratch_adjoint_equals_allocating_entry(hip, mode):
    from blackhorizon_amd.adjoint import Recorded, TraceSettings

    r_plus = 1.0 + math.sqrt(1.0 - SPIN * SPIN)
    # two calls in a row, the second smaller; the second uses a larger K
    # and the decreasing-step lane order
    for (w, h), K, sort in (((96, 72), 8, False), ((33, 17), 32, True)):
        cfg = TraceSettings(escape_radius=45.0, rtol=1e-8, atol=1e-10,
                            capture_margin=0.01 / r_plus, K=K, mode=mode, trajectory=False)
        s0 = _rays(w, h)
        n = s0.shape[0]
        print(f"static scratch adjoint: mode {mode}, {n} rays, K = {K}, sorted lanes {sort}",
              flush=True)
        rec = Recorded(s0, SPIN, cfg)
        g = torch.from_numpy(np.random.default_rng(n).normal(size=(n, 8))).cuda()
        order = torch.argsort(rec.ev_step, descending=True) if sort else None
        (gy_a, gs_a), (gy_s, gs_s) = _both(rec, g, order)
        status = rec.status.cpu().numpy()
        assert (status == 1).any() and (status == 2).any()
        assert np.all(np.isfinite(gy_a)) and np.all(np.isfinite(gs_a))
        np.testing.assert_array_equal(gy_s, gy_a)
        np.testing.assert_array_equal(gs_s, gs_a)
