"""GPU: the discrete-adjoint kernels against the FP64 PyTorch autograd
oracle (same discrete map: exact to rounding) and against central finite
differences of the whole adaptive pipeline (which also sees the step-size
controller react, an O(rtol)-level difference on escaped rays)."""

import math

import numpy as np
import pytest
import torch

from oracle import ad_oracle as ado
from oracle import bh_oracle as orc

pytestmark = pytest.mark.gpu

SPIN, INC, DIST, FOV = 0.7, 45.0, 40.0, 30.0


def _settings(mode="auto", **over):
    from blackhorizon_amd.adjoint import TraceSettings

    r_plus = 1.0 + math.sqrt(1.0 - SPIN * SPIN)
    kw = dict(escape_radius=45.0, rtol=1e-8, atol=1e-10, capture_margin=0.01 / r_plus,
              mode=mode, K=8)
    kw.update(over)
    return TraceSettings(**kw)


@pytest.mark.parametrize("spin", [0.7, -0.9, 0.0])
def test_rhs_vjp_matches_autograd(hip, spin):
    from blackhorizon_amd import _lib

    rng = np.random.default_rng(4)
    n = 512
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    pos = rng.uniform(2.5, 40.0, size=(n, 1)) * u
    d = rng.normal(size=(n, 3))
    s = orc.build_state(pos, orc.null_momentum_from_velocity(orc.Kerr(spin=spin), pos, d))
    v = rng.normal(size=(n, 8))
    st = torch.from_numpy(s).requires_grad_(True)
    a = torch.tensor(spin, dtype=torch.float64, requires_grad=True)
    f = ado.rhs(1.0, a, st)
    gz, ga = torch.autograd.grad((f * torch.from_numpy(v)).sum(), [st, a])
    # per-ray spin term: differentiate each ray's v.f separately
    ga_rows = torch.autograd.functional.jacobian(
        lambda aa: (ado.rhs(1.0, aa, torch.from_numpy(s)) * torch.from_numpy(v)).sum(-1), a)
    lib = _lib.require_device()
    sd = torch.from_numpy(s).cuda()
    vd = torch.from_numpy(v).cuda()
    zb = torch.empty_like(sd)
    ab = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.check(lib.bh_rhs_vjp_f64(1.0, spin, _lib.ptr(sd), _lib.ptr(vd), n, _lib.ptr(zb),
                                  _lib.ptr(ab), _lib.stream_handle()), "vjp")
    ref = gz.numpy()
    scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-300)
    assert np.max(np.abs(zb.cpu().numpy() - ref) / scale) < 1e-11
    np.testing.assert_allclose(ab.cpu().numpy(), ga_rows.numpy(), rtol=1e-10,
                               atol=1e-12 * np.abs(ga_rows.numpy()).max())
    assert abs(float(ab.sum()) - float(ga)) <= 1e-10 * max(1.0, abs(float(ga)))


def _small_camera(w=12, h=8):
    st = orc.Kerr(spin=SPIN)
    cam = orc.PinholeCamera.from_orbit(DIST, INC, fov_degrees=FOV, width=w, height=h)
    return orc.camera_state(st, cam)


@pytest.mark.parametrize("mode", ["precise", "fast"])
def test_adjoint_matches_autograd_oracle(hip, mode):
    from blackhorizon_amd import adjoint

    s0 = _small_camera()
    n = s0.shape[0]
    cfg = _settings(mode)
    rng = np.random.default_rng(9)
    gY = rng.normal(size=(n, 8))
    gY[:, 0] = 0.0
    # oracle (CPU torch autograd over the same discrete map)
    so = torch.from_numpy(s0).requires_grad_(True)
    ao = torch.tensor(SPIN, dtype=torch.float64, requires_grad=True)
    Yo, sto, _ = ado.trace_terminal(so, ao, escape_radius=cfg.escape_radius, rtol=cfg.rtol,
                                    atol=cfg.atol, capture_margin=cfg.capture_margin,
                                    score_factor=cfg.score_factor)
    go_s, go_a = torch.autograd.grad((Yo * torch.from_numpy(gY)).sum(), [so, ao])
    # HIP
    sg = torch.from_numpy(s0).cuda().requires_grad_(True)
    ag = torch.tensor(SPIN, dtype=torch.float64, device="cuda", requires_grad=True)
    Yg, stg = adjoint.trace_terminal(sg, ag, cfg)
    ((Yg * torch.from_numpy(gY).cuda()).sum()).backward()
    np.testing.assert_array_equal(stg.cpu().numpy(), sto.numpy())
    Yd = Yg.detach().cpu().numpy()
    np.testing.assert_allclose(Yd[:, 1:4], Yo.detach().numpy()[:, 1:4], rtol=1e-8, atol=1e-8)
    gs = sg.grad.cpu().numpy()
    ref = go_s.numpy()
    scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-12)
    err = np.abs(gs - ref) / scale
    assert np.max(err) < 1e-6, (np.max(err), np.argmax(np.max(err, axis=1)))
    assert abs(float(ag.grad) - float(go_a)) <= 1e-6 * abs(float(go_a))


def _fit_target():
    from blackhorizon_amd import adjoint

    cfg = _settings()
    with torch.no_grad():
        tgt = adjoint.camera_state(0.75, 50.0, DIST, FOV, 24, 24)
        Yt, st_t = adjoint.trace_terminal(tgt, torch.tensor(0.75, dtype=torch.float64,
                                                            device="cuda"),
                                          _settings(capture_margin=0.01 / (1 + math.sqrt(
                                              1 - 0.75 ** 2))))
    return adjoint.terminal_directions(Yt), st_t, cfg


def _loss(params, target, cfg, classes=(1, 2)):
    from blackhorizon_amd import adjoint

    a, inc, dist = params
    y0 = adjoint.camera_state(a, inc, dist, FOV, 24, 24)
    Y, st = adjoint.trace_terminal(y0, a, cfg)
    return adjoint.shadow_loss(Y, st, target[0], target[1], classes=classes)


@pytest.mark.parametrize("classes,tol", [((2,), 1e-5), ((1, 2), 3e-4)])
def test_shadow_fit_gradients_match_central_differences(hip, classes, tol):
    """d loss / d (spin, inclination, distance) by the adjoint vs central
    differences with h = 1e-6 over the full adaptive pipeline.  Escaped rays
    (direction on the escape sphere) meet 1e-5; with captured rays scored on
    the scoring sphere the controller's own response adds up to ~1e-4."""
    dirs, st_t, cfg = _fit_target()
    p = [torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True)
         for v in (SPIN, INC, DIST)]
    L = _loss(p, (dirs, st_t), cfg, classes)
    L.backward()
    grads = [float(t.grad) for t in p]
    h = 1e-6
    for k in range(3):
        with torch.no_grad():
            vals = [SPIN, INC, DIST]
            up = list(vals)
            dn = list(vals)
            up[k] += h
            dn[k] -= h
            lp = float(_loss([torch.tensor(v, dtype=torch.float64, device="cuda") for v in up],
                             (dirs, st_t), cfg, classes))
            lm = float(_loss([torch.tensor(v, dtype=torch.float64, device="cuda") for v in dn],
                             (dirs, st_t), cfg, classes))
        fd = (lp - lm) / (2 * h)
        print(f"param {k}: fd {fd:.12g} adjoint {grads[k]:.12g} rel {abs(fd - grads[k]) / abs(fd):.2e}")
        assert abs(fd - grads[k]) <= tol * abs(fd), (k, fd, grads[k])


def test_per_ray_momentum_gradients_match_central_differences(hip):
    from blackhorizon_amd import adjoint

    s0 = _small_camera(16, 4)
    cfg = _settings()
    rng = np.random.default_rng(5)
    w = torch.from_numpy(rng.normal(size=(s0.shape[0], 3))).cuda()
    sg = torch.from_numpy(s0).cuda().requires_grad_(True)
    Y, st = adjoint.trace_terminal(sg, SPIN, cfg)
    (adjoint.terminal_directions(Y) * w).sum().backward()
    g = sg.grad.cpu().numpy()
    esc = np.nonzero(st.cpu().numpy() == 2)[0][:6]
    assert len(esc) >= 3
    h = 1e-6
    for i in esc:
        for j in (5, 6, 7):
            d = np.zeros_like(s0)
            d[i, j] = h
            with torch.no_grad():
                Yp, _ = adjoint.trace_terminal(torch.from_numpy(s0 + d).cuda(), SPIN, cfg)
                Ym, _ = adjoint.trace_terminal(torch.from_numpy(s0 - d).cuda(), SPIN, cfg)
                fd = float(((adjoint.terminal_directions(Yp) - adjoint.terminal_directions(Ym))
                            * w).sum()) / (2 * h)
            assert abs(fd - g[i, j]) <= 1e-5 * abs(fd) + 1e-9, (i, j, fd, g[i, j])


def test_checkpoint_interval_does_not_change_gradients(hip):
    from blackhorizon_amd import adjoint

    s0 = _small_camera(8, 8)
    gY = torch.from_numpy(np.random.default_rng(2).normal(size=(64, 8))).cuda()
    out = []
    for K in (1, 5, 32, 4096):
        sg = torch.from_numpy(s0).cuda().requires_grad_(True)
        Y, _ = adjoint.trace_terminal(sg, SPIN, _settings(K=K))
        (Y * gY).sum().backward()
        out.append(sg.grad.cpu().numpy())
    for o in out[1:]:
        np.testing.assert_array_equal(o, out[0])


def test_backward_lane_order_does_not_change_gradients(hip):
    """Rays handed to lanes by decreasing step count (the default) or in
    index order give identical gradients, across a partial last wavefront."""
    from blackhorizon_amd import adjoint

    s0 = _small_camera(20, 7)  # 140 rays: two full wavefronts and a partial one
    gY = torch.from_numpy(np.random.default_rng(3).normal(size=(140, 8))).cuda()
    out = []
    for sort in (True, False):
        sg = torch.from_numpy(s0).cuda().requires_grad_(True)
        ag = torch.tensor(SPIN, dtype=torch.float64, device="cuda", requires_grad=True)
        Y, _ = adjoint.trace_terminal(sg, ag, _settings(sort_backward=sort))
        (Y * gY).sum().backward()
        out.append((sg.grad.cpu().numpy(), float(ag.grad)))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    # the spin gradient sums per-ray terms in a fixed order either way
    assert out[0][1] == out[1][1]
