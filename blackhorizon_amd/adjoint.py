"""Differentiable tracing: a discrete adjoint of the adaptive DP5(4) trace.

``trace_terminal(state0, spin, ...)`` is a ``torch.autograd.Function``: the
forward traces every ray exactly like ``trace_rays`` (same controller, same
statuses) with checkpoints every ``K`` accepted steps and returns each ray's
terminal point -- for escaped rays the point where the last step's cubic
Hermite dense output crosses the escape sphere, for captured rays where the
step that first enters the scoring sphere r = score_factor r_+ crosses it
(captured rays approach the horizon with diverging momenta, where a
fixed-step derivative is meaningless), otherwise the final state.  The
backward runs the HIP adjoint kernel: it pulls the terminal adjoint back
through the event, then through every accepted step with its step size held
fixed, replaying each checkpoint segment, and returns d/d(initial state) per
ray and d/d(spin).  The step-size controller itself
is not differentiated (the discrete adjoint of the accepted-step sequence),
so central differences, which also see the controller react, agree to
O(rtol).

The reference has no derivatives; this extends its tracer (tracer.py:50-157,
geodesics.py:42-74) for parameter fits such as spin and inclination from a
shadow image.  ``camera_state`` builds the camera rays with torch operations
so that inclination, observer distance and spin gradients flow into the
initial states (camera.py:48-113, geodesics.py:109-170).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import _lib
from .kerr import KerrSpacetime
from .tracer import make_params, trace_rays

TINY = 1e-30
CKPT_VALUES = 10              # state (8), controller step size, affine parameter
CKPT_BUDGET_BYTES = 16 << 30  # checkpoint memory reserved from max_steps alone


@dataclass
class TraceSettings:
    """trace_rays keyword arguments (tracer.py:50-61) plus the checkpoint
    interval K, the checkpoint capacity per ray (None: enough for
    max_steps), the scoring sphere of captured rays, r = score_factor r_+,
    an optional affine-parameter end, and whether the backward hands rays to
    wavefront lanes in order of decreasing step count (same results; fewer
    idle lanes)."""

    escape_radius: float
    max_steps: int = 20000
    rtol: float = 1e-9
    atol: float = 1e-12
    initial_step: float = 0.1
    max_step: float = 2.0
    min_step: float = 1e-10
    capture_margin: float = 1e-3
    mass: float = 1.0
    mode: str = "auto"
    K: int = 32
    cap: int | None = None
    score_factor: float = 3.0
    lambda_end: float = math.inf
    sort_backward: bool = True


class Recorded:
    """Device buffers of one recorded forward trace."""

    def __init__(self, state0: torch.Tensor, spin: float, cfg: TraceSettings):
        lib = _lib.require_device()
        dev = state0.device
        n = state0.shape[0]
        self.n = n
        self.cfg = cfg
        st = KerrSpacetime(cfg.mass, spin)
        self.params = make_params(st, cfg.escape_radius, cfg.max_steps, cfg.rtol, cfg.atol,
                                  cfg.initial_step, cfg.max_step, cfg.min_step,
                                  cfg.capture_margin, cfg.mode, cfg.lambda_end)
        self.score_radius = cfg.score_factor * st.outer_horizon_radius
        cap = cfg.cap if cfg.cap is not None else cfg.max_steps // cfg.K + 2
        if cfg.cap is None and cap * CKPT_VALUES * 8 * max(n, 1) > CKPT_BUDGET_BYTES:
            # max_steps would reserve more than the budget: size the
            # checkpoints from the step counts of a plain forward instead
            pre = trace_rays(st, state0, cfg.escape_radius, max_steps=cfg.max_steps,
                             rtol=cfg.rtol, atol=cfg.atol, initial_step=cfg.initial_step,
                             max_step=cfg.max_step, min_step=cfg.min_step,
                             capture_margin=cfg.capture_margin, mode=cfg.mode,
                             lambda_end=cfg.lambda_end)
            cap = int(pre.steps.max().item()) // cfg.K + 2 if n else 1
        self.cap = max(int(cap), 1)
        ld = max(n, 1)
        f64 = dict(dtype=torch.float64, device=dev)
        stream = _lib.stream_handle()
        self.soa_in = torch.empty((8, ld), **f64)
        _lib.check(lib.bh_aos_to_soa_f64(_lib.ptr(state0.contiguous()), _lib.ptr(self.soa_in), ld,
                                         n, stream), "aos_to_soa")
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)
        _lib.check(lib.bh_count_outside_f64(self.params, _lib.ptr(self.soa_in), ld, n,
                                            _lib.ptr(cnt), stream), "count_outside")
        if int(cnt.item()) > 0:
            raise ValueError("escape_radius must exceed all initial radii")
        self.soa_out = torch.empty_like(self.soa_in)
        self.status = torch.empty(n, dtype=torch.int32, device=dev)
        self.steps = torch.empty(n, dtype=torch.int64, device=dev)
        self.attempts = torch.empty(n, dtype=torch.int64, device=dev)
        self.iters = torch.zeros(1, dtype=torch.int64, device=dev)
        self.ckpt = torch.empty((self.cap, CKPT_VALUES, ld), **f64)
        self.n_ckpt = torch.empty(n, dtype=torch.int32, device=dev)
        self.term = torch.empty((8, ld), **f64)
        self.prev = torch.empty((9, ld), **f64)
        self.evy1 = torch.empty((8, ld), **f64)
        self.ev_step = torch.empty(n, dtype=torch.int64, device=dev)
        self.theta = torch.empty(n, **f64)
        self.event = torch.empty(n, dtype=torch.int32, device=dev)
        _lib.check(lib.bh_trace_record_f64(
            self.params, _lib.ptr(self.soa_in), ld, _lib.ptr(self.soa_out), ld,
            _lib.ptr(self.status), _lib.ptr(self.steps), _lib.ptr(self.attempts),
            _lib.ptr(self.iters), n, cfg.K, self.cap, self.score_radius, _lib.ptr(self.ckpt),
            _lib.ptr(self.n_ckpt), _lib.ptr(self.term), _lib.ptr(self.prev),
            _lib.ptr(self.evy1), _lib.ptr(self.ev_step), _lib.ptr(self.theta),
            _lib.ptr(self.event), stream), "trace_record")

    def terminal(self) -> torch.Tensor:
        return self.term[:, : self.n].t().contiguous()

    def final_states(self) -> torch.Tensor:
        return self.soa_out[:, : self.n].t().contiguous()

    def backward(self, g_terminal: torch.Tensor):
        """(d/d state0 (n, 8), per-ray d/d spin (n,)) for a terminal adjoint (n, 8)."""
        lib = _lib.require_device()
        n = self.n
        ld = max(n, 1)
        dev = self.term.device
        if int((self.event < 0).sum().item()) > 0:
            raise RuntimeError("a ray exceeded the checkpoint capacity; raise cap")
        gterm = torch.zeros((8, ld), dtype=torch.float64, device=dev)
        gterm[:, :n] = g_terminal.t()
        gy0 = torch.empty((8, ld), dtype=torch.float64, device=dev)
        gspin = torch.empty(ld, dtype=torch.float64, device=dev)
        per_ray = lib.bh_adjoint_scratch_doubles(self.cfg.K)
        if per_ray < 0:
            raise ValueError("checkpoint interval K out of range")
        # one block of per_ray x 64 doubles per wavefront of 64 rays
        scratch = torch.empty(per_ray * (-(-n // 64) * 64 or 64), dtype=torch.float64,
                              device=dev)
        # longest reversals first, rays of similar length in one wavefront
        order = torch.argsort(self.ev_step, descending=True) if self.cfg.sort_backward and n \
            else None
        _lib.check(lib.bh_trace_adjoint_f64(
            self.params, n, self.cfg.K, self.score_radius, _lib.ptr(self.ckpt),
            _lib.ptr(self.event), _lib.ptr(self.ev_step), _lib.ptr(self.evy1),
            _lib.ptr(self.prev), _lib.ptr(self.theta), _lib.ptr(gterm), _lib.ptr(gy0),
            _lib.ptr(gspin), _lib.ptr(scratch), _lib.ptr(order), _lib.stream_handle()),
            "trace_adjoint")
        return gy0[:, :n].t().contiguous(), gspin[:n]


class _TraceTerminal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0, spin, cfg):
        rec = Recorded(state0.detach().to(torch.float64).contiguous(), float(spin), cfg)
        ctx.rec = rec
        ctx.spin_shape = spin.shape
        status = rec.status.clone()
        ctx.mark_non_differentiable(status)
        return rec.terminal(), status

    @staticmethod
    def backward(ctx, g_term, g_status):
        gy0, gspin = ctx.rec.backward(g_term.to(torch.float64))
        return gy0, gspin.sum().reshape(ctx.spin_shape), None


def trace_terminal(state0: torch.Tensor, spin: torch.Tensor, cfg: TraceSettings):
    """Differentiable trace: (terminal points (n, 8), status (n,))."""
    if not isinstance(spin, torch.Tensor):
        spin = torch.tensor(float(spin), dtype=torch.float64, device=state0.device)
    return _TraceTerminal.apply(state0, spin, cfg)


# ---------------------------------------------------------------------------
# differentiable camera rays (torch operations, float64)
# ---------------------------------------------------------------------------

def _ks_h_l(mass, spin, x, y, z):
    """Kerr-Schild H and l at points (kerr.py:136-152), torch-differentiable."""
    a2 = spin * spin
    rho2 = x * x + y * y + z * z
    q = rho2 - a2
    s = torch.clamp(torch.sqrt(q * q + 4.0 * a2 * z * z), min=TINY)
    r2 = 0.5 * (q + s)
    r = torch.sqrt(torch.clamp(r2, min=TINY))
    qd = torch.clamp(r2 * r2 + a2 * z * z, min=TINY)
    h = mass * r * r2 / qd
    inv = 1.0 / torch.clamp(r2 + a2, min=TINY)
    l = torch.stack([(r * x + spin * y) * inv, (r * y - spin * x) * inv, z / r], dim=-1)
    return h, l


def null_momentum_torch(mass, spin, positions, directions):
    """Past-directed, |p_t| = 1 photon covectors (geodesics.py:109-170)."""
    h, l = _ks_h_l(mass, spin, positions[:, 0], positions[:, 1], positions[:, 2])
    l_dot_s = (l * directions).sum(-1)
    s_dot_s = (directions * directions).sum(-1)
    g_tt = 2.0 * h - 1.0
    b = 4.0 * h * l_dot_s
    c = s_dot_s + 2.0 * h * l_dot_s * l_dot_s
    sq = torch.sqrt(torch.clamp(b * b - 4.0 * g_tt * c, min=0.0))
    k_t = torch.minimum((-b + sq) / (2.0 * g_tt), (-b - sq) / (2.0 * g_tt))
    l_dot_k = k_t + l_dot_s
    m0 = -k_t + 2.0 * h * l_dot_k
    mom = torch.cat([m0[:, None], directions + (2.0 * h * l_dot_k)[:, None] * l], dim=-1)
    return mom / mom[:, :1].abs()


def camera_state(spin, inclination_degrees, distance, fov_degrees: float, width: int,
                 height: int, mass: float = 1.0, azimuth_degrees: float = 0.0,
                 device=None) -> torch.Tensor:
    """(width * height, 8) past-directed camera rays, differentiable in spin,
    inclination and distance (tensors or floats); pixel order and basis as
    PinholeCamera.from_orbit(...).ray_directions() (camera.py:48-113)."""
    dev = device or torch.device("cuda", torch.cuda.current_device())
    t = lambda v: v if isinstance(v, torch.Tensor) else torch.tensor(float(v),  # noqa: E731
                                                                     dtype=torch.float64,
                                                                     device=dev)
    spin, inc_deg, dist = t(spin), t(inclination_degrees), t(distance)
    inc = inc_deg * (math.pi / 180.0)
    azi = math.radians(azimuth_degrees)
    pos = torch.stack([dist * torch.sin(inc) * math.cos(azi), dist * torch.sin(inc) * math.sin(azi),
                       dist * torch.cos(inc)])
    fwd = -pos / torch.linalg.norm(pos)
    up_hint = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64, device=dev)
    right = torch.linalg.cross(fwd, up_hint)
    right = right / torch.linalg.norm(right)
    up = torch.linalg.cross(right, fwd)
    half_w = math.tan(math.radians(fov_degrees) / 2.0)
    half_h = half_w * height / width
    cols = (torch.arange(width, dtype=torch.float64, device=dev) + 0.5) / width * 2.0 - 1.0
    rows = 1.0 - (torch.arange(height, dtype=torch.float64, device=dev) + 0.5) / height * 2.0
    v, u = torch.meshgrid(rows * half_h, cols * half_w, indexing="ij")
    d = fwd[None, None, :] + u[..., None] * right[None, None, :] + v[..., None] * up[None, None, :]
    d = (d / torch.linalg.norm(d, dim=-1, keepdim=True)).reshape(-1, 3)
    p = pos[None, :].expand(d.shape[0], 3)
    mom = null_momentum_torch(mass, spin, p, d)
    t0 = torch.zeros((d.shape[0], 1), dtype=torch.float64, device=dev)
    return torch.cat([t0, p, mom], dim=-1)


# ---------------------------------------------------------------------------
# shadow-fit loss
# ---------------------------------------------------------------------------

def shadow_loss(terminal: torch.Tensor, status: torch.Tensor, target_dirs: torch.Tensor,
                target_status: torch.Tensor, mismatch_cost: float = 1.0,
                classes: tuple = (1, 2)) -> torch.Tensor:
    """Sum over rays of |n - n_target|^2 with n the unit position of the
    terminal point (the squared chord between the two directions, the smooth
    form of the squared angular distance), for rays whose class matches the
    target's: escaped rays by where they leave through the escape sphere,
    captured rays by where they plunge through the scoring sphere near the
    horizon; rays whose class differs cost ``mismatch_cost`` each.
    ``classes`` limits the scored classes (1 captured, 2 escaped)."""
    pos = terminal[:, 1:4]
    n = pos / torch.linalg.norm(pos, dim=-1, keepdim=True)
    scored = torch.zeros_like(status, dtype=torch.bool)
    for c in classes:
        scored |= status == c
    same = (status == target_status) & scored
    per_ray = ((n - target_dirs) ** 2).sum(-1)
    return torch.where(same, per_ray, torch.full_like(per_ray, mismatch_cost)).sum()


def terminal_directions(terminal: torch.Tensor) -> torch.Tensor:
    pos = terminal[:, 1:4]
    return pos / torch.linalg.norm(pos, dim=-1, keepdim=True)
