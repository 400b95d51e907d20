"""Derivative oracle: the reference's adaptive DP5(4) trace in FP64 PyTorch,
differentiated by torch.autograd.

TEST INFRASTRUCTURE ONLY (see oracle/bh_oracle.py for the import rule).

It restates, with torch tensor operations on the CPU, the reference's
geometry (kerr.py:118-193), right-hand side (geodesics.py:42-74), DP5(4)
step, error ratio and step factor (integrators.py:76-124) and the per-ray
controller of trace_rays (tracer.py:101-154), vectorised over rays with
masks instead of the active-set gather.  Step sizes and accept/reject
decisions are taken on detached values, so autograd differentiates the
accepted-step sequence with its step sizes held fixed -- the quantity the
HIP adjoint kernel computes.  The terminal point of an escaped ray is where
the cubic Hermite dense output of its last step,
Y(theta) = h00 y0 + h10 h f0 + h01 y1 + h11 h f1, crosses the escape
sphere; a captured ray's is where the step that first enters the scoring
sphere r = score_factor * r_+ crosses it; theta is found by bisection on
detached values and
re-attached to the graph by one Newton step, which carries the implicit
derivative d theta = -dg / g'.

Pinned by central finite differences of its own terminal points
(tests/test_oracle_adjoint.py) and by the numpy oracle's statuses and step
counts on the same rays.
"""

from __future__ import annotations

import torch

TINY = 1e-30

A21 = 1.0 / 5.0
A31, A32 = 3.0 / 40.0, 9.0 / 40.0
A41, A42, A43 = 44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0
A51, A52, A53, A54 = 19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0
A61, A62, A63, A64, A65 = (9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0,
                           49.0 / 176.0, -5103.0 / 18656.0)
B1, B3, B4, B5, B6 = 35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0
E1, E3, E4, E5, E6, E7 = (71.0 / 57600.0, -71.0 / 16695.0, 71.0 / 1920.0,
                          -17253.0 / 339200.0, 22.0 / 525.0, -1.0 / 40.0)


def geometry(m, a, x, y, z):
    """r, H, l, grad H, grad l (kerr.py:136-189); a may be a tensor."""
    a2 = a * a
    rho2 = x * x + y * y + z * z
    q = rho2 - a2
    s = torch.clamp(torch.sqrt(q * q + 4.0 * a2 * z * z), min=TINY)
    r2 = 0.5 * (q + s)
    r = torch.sqrt(torch.clamp(r2, min=TINY))
    r2a2 = r2 + a2
    qd = torch.clamp(r2 * r2 + a2 * z * z, min=TINY)
    h = m * r * r2 / qd
    inv_w = 1.0 / torch.clamp(r2a2, min=TINY)
    inv_r = 1.0 / r
    lx = (r * x + a * y) * inv_w
    ly = (r * y - a * x) * inv_w
    lz = z * inv_r
    inv_s = 1.0 / s
    r_x = x * r * inv_s
    r_y = y * r * inv_s
    r_z = z * r2a2 * inv_r * inv_s
    inv_qd2 = 1.0 / (qd * qd)
    dh_dr = m * r2 * (3.0 * a2 * z * z - r2 * r2) * inv_qd2
    dh_dz = -2.0 * m * a2 * z * r * r2 * inv_qd2
    gh = torch.stack([dh_dr * r_x, dh_dr * r_y, dh_dr * r_z + dh_dz], -1)
    two_r = 2.0 * r
    gl = torch.stack([
        torch.stack([(r_x * x + r) * inv_w - lx * two_r * r_x * inv_w,
                     (r_x * y - a) * inv_w - ly * two_r * r_x * inv_w,
                     -lz * r_x * inv_r], -1),
        torch.stack([(r_y * x + a) * inv_w - lx * two_r * r_y * inv_w,
                     (r_y * y + r) * inv_w - ly * two_r * r_y * inv_w,
                     -lz * r_y * inv_r], -1),
        torch.stack([r_z * x * inv_w - lx * two_r * r_z * inv_w,
                     r_z * y * inv_w - ly * two_r * r_z * inv_w,
                     (1.0 - lz * r_z) * inv_r], -1),
    ], -2)
    return r, h, torch.stack([lx, ly, lz], -1), gh, gl


def rhs(m, a, state):
    """geodesic_rhs (geodesics.py:42-74)."""
    r, h, l, gh, gl = geometry(m, a, state[:, 1], state[:, 2], state[:, 3])
    p_t = state[:, 4]
    ps = state[:, 5:8]
    lp = -p_t + (l * ps).sum(-1)
    two_h_lp = 2.0 * h * lp
    dt = -p_t + two_h_lp
    dx = ps - two_h_lp[:, None] * l
    dp = gh * (lp * lp)[:, None] + two_h_lp[:, None] * (gl * ps[:, None, :]).sum(-1)
    return torch.cat([dt[:, None], dx, torch.zeros_like(p_t)[:, None], dp], -1)


def radius(a, pos):
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    a2 = a * a
    q = x * x + y * y + z * z - a2
    s = torch.clamp(torch.sqrt(q * q + 4.0 * a2 * z * z), min=TINY)
    return torch.sqrt(torch.clamp(0.5 * (q + s), min=TINY))


def dp5(m, a, y, h):
    """dormand_prince_step (integrators.py:76-107)."""
    hh = h[:, None]
    k1 = rhs(m, a, y)
    k2 = rhs(m, a, y + hh * (A21 * k1))
    k3 = rhs(m, a, y + hh * (A31 * k1 + A32 * k2))
    k4 = rhs(m, a, y + hh * (A41 * k1 + A42 * k2 + A43 * k3))
    k5 = rhs(m, a, y + hh * (A51 * k1 + A52 * k2 + A53 * k3 + A54 * k4))
    k6 = rhs(m, a, y + hh * (A61 * k1 + A62 * k2 + A63 * k3 + A64 * k4 + A65 * k5))
    new = y + hh * (B1 * k1 + B3 * k3 + B4 * k4 + B5 * k5 + B6 * k6)
    k7 = rhs(m, a, new)
    err = hh * (E1 * k1 + E3 * k3 + E4 * k4 + E5 * k5 + E6 * k6 + E7 * k7)
    return new, err


def error_ratio(y, new, err, rtol, atol):
    """integrators.py:110-117 with NumPy's pairwise order for the mean of 8."""
    sc = atol + rtol * torch.maximum(y.abs(), new.abs())
    q = (err / sc) ** 2
    tot = ((q[:, 0] + q[:, 1]) + (q[:, 2] + q[:, 3])) + ((q[:, 4] + q[:, 5]) + (q[:, 6] + q[:, 7]))
    ratio = torch.sqrt(tot / 8.0)
    return torch.where(torch.isfinite(ratio), ratio, torch.full_like(ratio, float("inf")))


def step_factor(ratio):
    """integrators.py:120-124."""
    return torch.clamp(0.9 * torch.clamp(ratio, min=1e-16) ** (-0.2), 0.2, 5.0)


def _hermite(th, h, y0, f0, y1, f1, deriv=False):
    t2, t3 = th * th, th * th * th
    if deriv:
        b = (6 * t2 - 6 * th, 3 * t2 - 4 * th + 1, -6 * t2 + 6 * th, 3 * t2 - 2 * th)
    else:
        b = (2 * t3 - 3 * t2 + 1, t3 - 2 * t2 + th, -2 * t3 + 3 * t2, t3 - t2)
    return (b[0][:, None] * y0 + (b[1] * h)[:, None] * f0 + b[2][:, None] * y1
            + (b[3] * h)[:, None] * f1)


def trace_terminal(state0, spin, escape_radius, max_steps=20000, rtol=1e-9, atol=1e-12,
                   initial_step=0.1, max_step=2.0, min_step=1e-10, capture_margin=1e-3,
                   mass=1.0, score_factor=3.0, return_steps=False, lambda_end=float("inf")):
    """Terminal points (n, 8) and statuses of a trace, differentiable in
    ``state0`` and ``spin`` through the accepted steps (fixed step sizes).

    Escaped rays end where their last step crosses the escape sphere.
    Captured rays are scored where they first cross the sphere
    r = score_factor * r_+ inward: past-directed rays in these
    (ingoing) coordinates approach the horizon with diverging momenta and
    shrinking steps, where a derivative of the state with the step sizes held
    fixed no longer approximates the true one, while the crossing of a
    sphere outside the horizon is well conditioned."""
    n = state0.shape[0]
    a_val = float(spin.detach()) if isinstance(spin, torch.Tensor) else float(spin)
    capture_radius = (mass + (mass * mass - a_val * a_val) ** 0.5) * (1.0 + capture_margin)
    y = state0
    dummy = state0.detach().clone()
    h = torch.full((n,), float(initial_step), dtype=torch.float64)
    steps = torch.zeros(n, dtype=torch.int64)
    status = torch.zeros(n, dtype=torch.int32)
    y_prev = state0
    h_prev = torch.zeros(n, dtype=torch.float64)
    r_score = (mass + (mass * mass - a_val * a_val) ** 0.5) * score_factor
    crossed = torch.zeros(n, dtype=torch.bool)
    ys0, ys1, hs = state0, state0, torch.zeros(n, dtype=torch.float64)
    r_cur = radius(a_val, state0.detach()[:, 1:4])
    lam = torch.zeros(n, dtype=torch.float64)
    h_hist = []
    for _ in range(4 * max_steps):
        active = status == 0
        if not bool(active.any()):
            break
        y_in = torch.where(active[:, None], y, dummy)
        ht = torch.minimum(h, lambda_end - lam)  # integrate_dense's clip to the end
        new, err = dp5(mass, spin, y_in, ht)
        ratio = error_ratio(y_in.detach(), new.detach(), err.detach(), rtol, atol)
        accept = (ratio <= 1.0) & active
        y_prev = torch.where(accept[:, None], y, y_prev)
        h_prev = torch.where(accept, ht, h_prev)
        h_hist.append(torch.where(accept, ht, torch.full_like(ht, float("nan"))))
        y = torch.where(accept[:, None], new, y)
        steps = steps + accept.to(torch.int64)
        lam = torch.where(accept, lam + ht, lam)
        h = torch.where(active, torch.clamp(ht * step_factor(ratio), 0.0, max_step), h)
        yd = y.detach()
        rad = radius(a_val, yd[:, 1:4])
        first = accept & ~crossed & (r_cur > r_score) & (rad <= r_score)
        ys0 = torch.where(first[:, None], y_prev, ys0)
        ys1 = torch.where(first[:, None], y, ys1)
        hs = torch.where(first, h_prev, hs)
        crossed = crossed | first
        r_cur = torch.where(accept, rad, r_cur)
        fin = torch.isfinite(yd).all(-1)
        s = status.clone()
        s = torch.where(active & ~fin, torch.full_like(s, 4), s)
        s = torch.where(active & fin & (rad <= capture_radius), torch.full_like(s, 1), s)
        s = torch.where(active & fin & (rad >= escape_radius), torch.full_like(s, 2), s)
        still = active & (s == 0)
        s = torch.where(still & (h < min_step), torch.full_like(s, 4), s)
        s = torch.where(still & (steps >= max_steps), torch.full_like(s, 3), s)
        s = torch.where((s == 0) & (lam >= lambda_end), torch.full_like(s, 6), s)
        status = s
    status = torch.where(status == 0, torch.full_like(status, 3), status)
    esc = (status == 2) & (steps > 0)
    cap = (status == 1) & crossed
    ev = esc | cap
    extra = ()
    if return_steps:
        hh = torch.stack(h_hist, 1) if h_hist else torch.zeros((n, 0), dtype=torch.float64)
        extra = ([hh[i][~torch.isnan(hh[i])] for i in range(n)],)
    if not bool(ev.any()):
        return (y, status, steps) + extra
    idx = torch.nonzero(ev)[:, 0]
    is_esc = esc[idx]
    y0 = torch.where(is_esc[:, None], y_prev[idx], ys0[idx])
    y1 = torch.where(is_esc[:, None], y[idx], ys1[idx])
    hp = torch.where(is_esc, h_prev[idx], hs[idx])
    r_plus_t = mass + torch.sqrt(mass * mass - spin * spin)
    Rr = torch.where(is_esc, torch.full_like(hp, float(escape_radius)), r_plus_t * score_factor)
    term = _event_point(mass, spin, a_val, y0, y1, hp, Rr)
    out = y.clone()
    out = out.index_put((idx,), term)
    return (out, status, steps) + extra


def _event_point(mass, spin, a_val, y0, y1, hp, Rr):
    """Point where the Hermite dense output of the steps y0 -> y1 (sizes hp)
    crosses r = Rr, differentiable with the implicit derivative of theta."""
    f0, f1 = rhs(mass, spin, y0), rhs(mass, spin, y1)
    with torch.no_grad():
        Rd = Rr.detach()
        g = lambda t: radius(a_val, _hermite(t, hp, y0, f0, y1, f1)[:, 1:4]) - Rd  # noqa: E731
        lo = torch.zeros(y0.shape[0], dtype=torch.float64)
        hi = torch.ones(y0.shape[0], dtype=torch.float64)
        sgn0 = torch.sign(g(lo))
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            same = torch.sign(g(mid)) == sgn0
            lo = torch.where(same, mid, lo)
            hi = torch.where(same, hi, mid)
        th = 0.5 * (lo + hi)
    # one Newton step with the graph: theta = th - g(th) / g'(th)
    Yt = _hermite(th, hp, y0, f0, y1, f1)
    dYt = _hermite(th, hp, y0, f0, y1, f1, deriv=True)
    rr = radius(spin, Yt[:, 1:4])
    pos = Yt[:, 1:4].detach().requires_grad_(True)
    with torch.enable_grad():
        gr = torch.autograd.grad(radius(a_val, pos).sum(), pos)[0]
    gprime = (gr * dYt[:, 1:4].detach()).sum(-1)
    th_ad = th - (rr - Rr) / gprime
    return _hermite(th_ad, hp, y0, f0, y1, f1)


def terminal_fixed_steps(state0, spin, h_lists, escape_radius, capture_margin=1e-3, mass=1.0,
                         score_factor=3.0):
    """The terminal points with every ray's accepted step sizes given
    (``h_lists[i]``: its accepted steps' sizes, in order): exactly the map
    the discrete adjoint differentiates.  Central differences of this
    function pin the adjoint without the controller's response."""
    a_val = float(spin.detach()) if isinstance(spin, torch.Tensor) else float(spin)
    r_plus = mass + (mass * mass - a_val * a_val) ** 0.5
    capture_radius = r_plus * (1.0 + capture_margin)
    r_score = r_plus * score_factor
    rows = []
    for i in range(state0.shape[0]):
        y = state0[i:i + 1]
        r_prev = float(radius(a_val, y.detach()[:, 1:4]))
        score_ev = last_ev = None
        for h in h_lists[i].tolist():
            new, _ = dp5(mass, spin, y, torch.tensor([h], dtype=torch.float64))
            r_new = float(radius(a_val, new.detach()[:, 1:4]))
            if score_ev is None and r_prev > r_score and r_new <= r_score:
                score_ev = (y, new, h)
            last_ev = (y, new, h)
            y, r_prev = new, r_new
        if last_ev is not None and r_prev >= escape_radius:
            y0, y1, h = last_ev
            R = torch.tensor([float(escape_radius)], dtype=torch.float64)
        elif r_prev <= capture_radius and score_ev is not None:
            y0, y1, h = score_ev
            sp = spin if isinstance(spin, torch.Tensor) else torch.tensor(float(spin),
                                                                          dtype=torch.float64)
            R = ((mass + torch.sqrt(mass * mass - sp * sp)) * score_factor).reshape(1)
        else:
            rows.append(y)
            continue
        rows.append(_event_point(mass, spin, a_val, y0, y1,
                                 torch.tensor([h], dtype=torch.float64), R))
    return torch.cat(rows, 0)
