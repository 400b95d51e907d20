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

import dataclasses
import math
from dataclasses import dataclass

import torch

from . import _lib
from .kerr import KerrSpacetime
from .tracer import Tune, make_params, resolve_mode, trace_rays

TINY = 1e-30
CKPT_VALUES = 10              # state (8), controller step size, affine parameter
CKPT_BUDGET_BYTES = 16 << 30  # checkpoint memory reserved from max_steps alone
TRAJ_VALUES = 9               # per stored step: start state (8) and step size
TRAJ_PAGE_STEPS = 32          # accepted steps per page of the store (BH_TRAJ_PAGE_STEPS)
TRAJ_RUN_MAX = 16             # longest run of pages a ray takes at once (BH_TRAJ_RUN_MAX)
TRAJ_BUDGET_BYTES = 96 << 30  # largest trajectory store chosen automatically

# Pages of the trajectory store the last reversed recording of each
# configuration (ray count and trace settings other than the spin) used: a
# fit traces the same camera again and again, so the next pool is sized
# from this rather than from the (generous) expected step counts.  At most
# _POOL_HINTS_MAX configurations are kept; the least recently stored go first.
_POOL_HINTS: dict = {}
_POOL_HINTS_MAX = 64


def _remember_pages(key, pages: int) -> None:
    _POOL_HINTS.pop(key, None)
    _POOL_HINTS[key] = int(pages)
    while len(_POOL_HINTS) > _POOL_HINTS_MAX:
        del _POOL_HINTS[next(iter(_POOL_HINTS))]


@dataclass
class TraceSettings:
    """trace_rays keyword arguments (tracer.py:50-61) plus the checkpoint
    interval K, the checkpoint capacity per ray (None: enough for
    max_steps), the scoring sphere of captured rays, r = score_factor r_+,
    an optional affine-parameter end, and whether the recording forward and
    the backward hand rays to wavefront lanes in order of decreasing
    (expected) step count (same results; rays of similar length share a
    wavefront and finish together).

    ``trajectory``: store every accepted step (72 bytes each, in pages of
    TRAJ_PAGE_STEPS steps that rays take from a pool as they go) instead of
    checkpoints, so the backward reverses the stored steps without replaying
    them -- same gradients, bit for bit.  The pool is sized from the rays'
    expected step counts; should the recording find it too small, the
    exact per-ray counts of a plain trace size it and the recording runs
    again (at the backward: the forward does not wait for the device).  None
    chooses the store when the estimated pool fits TRAJ_BUDGET_BYTES and a
    quarter of the device memory; True always; False never (checkpoints)."""

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
    sort_forward: bool = True
    trajectory: bool | None = None


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
        order = None
        pre = None
        self.pool = self.page_link = self.last_page = None
        self.pool_pages = 0
        pool_pages = None
        est = self._expected_steps(st, state0, cfg) if n else None
        self._hint_key = (n, str(dev), cfg.escape_radius, cfg.max_steps, cfg.rtol, cfg.atol,
                          cfg.initial_step, cfg.max_step, cfg.min_step, cfg.capture_margin,
                          cfg.mass, resolve_mode(cfg.mode, cfg.rtol), cfg.lambda_end)
        if cfg.trajectory is not False:
            pages = self._pages(est, n)
            hint = _POOL_HINTS.get(self._hint_key)
            if hint is not None:
                # a quarter more than last time, and a page in 16 rays
                pages = min(pages, hint + hint // 4 + n // 16 + 64)
            if cfg.trajectory or self._store_fits(pages, dev):
                pool_pages = pages
        if pool_pages is not None and not self._store_fits(pool_pages, dev):
            # a store is asked for beyond the estimate's budget: exact sizes
            pre = self._pre_pass(st, state0, cfg)
            pool_pages = self._pages(pre.steps, n)
        elif pool_pages is None and cfg.cap is None and \
                cap * CKPT_VALUES * 8 * max(n, 1) > CKPT_BUDGET_BYTES:
            # max_steps would reserve more checkpoints than the budget: size
            # them from the step counts of a forward pass instead
            pre = self._pre_pass(st, state0, cfg)
            cap = int(pre.steps.max().item()) // cfg.K + 2 if n else 1
        if cfg.sort_forward and n:
            # longest rays first: the pre-pass's exact attempt counts, else,
            # for rays traced to an affine end, their expected steps
            if pre is not None:
                order = torch.argsort(pre.attempts, descending=True)
            elif math.isfinite(cfg.lambda_end):
                order = torch.from_numpy(est).to(dev).argsort(descending=True)
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
        self.term = torch.empty((8, ld), **f64)
        self.prev = torch.empty((9, ld), **f64)
        self.evy1 = torch.empty((8, ld), **f64)
        self.ev_step = torch.empty(n, dtype=torch.int64, device=dev)
        self.theta = torch.empty(n, **f64)
        self.event = torch.empty(n, dtype=torch.int32, device=dev)
        self.workspace = torch.zeros(2, dtype=torch.int64, device=dev)
        if pool_pages is not None:
            self.ckpt = self.n_ckpt = None
            # no host synchronisation here: whether the pool sufficed is
            # read by the backward (the terminal points, statuses and step
            # counts are right either way; only the store would be cut)
            self._redo = (st, state0)
            self._record_store(lib, pool_pages, order, stream)
            return
        self.ckpt = torch.empty((self.cap, CKPT_VALUES, ld), **f64)
        self.n_ckpt = torch.empty(n, dtype=torch.int32, device=dev)
        _lib.check(lib.bh_trace_record_f64(
            self.params, _lib.ptr(self.soa_in), ld, _lib.ptr(self.soa_out), ld,
            _lib.ptr(self.status), _lib.ptr(self.steps), _lib.ptr(self.attempts),
            _lib.ptr(self.iters), _lib.ptr(order), n, cfg.K, self.cap, self.score_radius,
            _lib.ptr(self.ckpt),
            _lib.ptr(self.n_ckpt), _lib.ptr(self.term), _lib.ptr(self.prev),
            _lib.ptr(self.evy1), _lib.ptr(self.ev_step), _lib.ptr(self.theta),
            _lib.ptr(self.event), _lib.ptr(self.workspace), stream), "trace_record")

    def _record_store(self, lib, pool_pages, order, stream):
        """The recording forward into a trajectory store of pool_pages pages."""
        n = self.n
        ld = max(n, 1)
        dev = self.term.device
        self.pool_pages = int(pool_pages)
        self.pool = self.page_link = None  # release a smaller store first
        self.pool = torch.empty(max(self.pool_pages, 1) * TRAJ_PAGE_STEPS * TRAJ_VALUES,
                                dtype=torch.float64, device=dev)
        self.page_link = torch.empty(max(self.pool_pages, 1), dtype=torch.int64, device=dev)
        self.last_page = torch.empty(n, dtype=torch.int64, device=dev)
        _lib.check(lib.bh_trace_record_trajectory_f64(
            self.params, _lib.ptr(self.soa_in), ld, _lib.ptr(self.soa_out), ld,
            _lib.ptr(self.status), _lib.ptr(self.steps), _lib.ptr(self.attempts),
            _lib.ptr(self.iters), _lib.ptr(order), n, _lib.ptr(self.pool), self.pool_pages,
            _lib.ptr(self.page_link), _lib.ptr(self.last_page), self.score_radius,
            _lib.ptr(self.term), _lib.ptr(self.prev), _lib.ptr(self.evy1),
            _lib.ptr(self.ev_step), _lib.ptr(self.theta), _lib.ptr(self.event),
            _lib.ptr(self.workspace), stream), "trace_record_trajectory")

    @staticmethod
    def _expected_steps(st, state0, cfg):
        """Expected accepted steps, at most max_steps: for rays integrated
        to an affine end batch.expected_steps, one per ray (host array); for
        rays traced to escape one value, escape_radius / max_step +
        3 rtol^-1/5 (half the near-hole term batch.expected_steps ranks
        with: camera rays take 88 steps on average at rtol 1e-8 where this
        gives 142 and the ranking estimate 261)."""
        import numpy as np

        from .batch import TraceJob, expected_steps

        if math.isfinite(cfg.lambda_end):
            est = expected_steps(TraceJob(st, state0, cfg.escape_radius, rtol=cfg.rtol,
                                          max_step=cfg.max_step, lambda_end=cfg.lambda_end))
        else:
            est = cfg.escape_radius / cfg.max_step + 3.0 * cfg.rtol ** -0.2
        cap = max(cfg.max_steps, 0)
        return np.minimum(est, cap) if np.ndim(est) else min(float(est), float(cap))

    @staticmethod
    def _pages(steps, n):
        """Pages the rays request for `steps` accepted steps each (a tensor
        or array per ray, or one number for all; at least one in all): a
        ray needing p pages takes runs of 1, 2, 4, ... TRAJ_RUN_MAX pages
        until they hold p."""
        if not n:
            return 1
        import numpy as np

        if isinstance(steps, torch.Tensor):
            # on the device: the same count in torch operations
            p = torch.div(steps.to(torch.int64) + TRAJ_PAGE_STEPS - 1, TRAJ_PAGE_STEPS,
                          rounding_mode="floor")
            total = torch.zeros_like(p)
            full, run = 0, 1
            while run < TRAJ_RUN_MAX:
                full += run
                total = torch.where((total == 0) & (p <= full), full, total)
                run *= 2
            rest = torch.div(torch.clamp(p - full, min=0) + TRAJ_RUN_MAX - 1, TRAJ_RUN_MAX,
                             rounding_mode="floor")
            total = torch.where(total > 0, total, full + TRAJ_RUN_MAX * rest)
            total = torch.where(p <= 0, 0, total)
            return max(int(total.sum().item()), 1)
        p = np.ceil(np.asarray(steps, dtype=np.float64) / TRAJ_PAGE_STEPS)
        short = np.zeros_like(p)
        full = 0.0
        run = 1
        while run < TRAJ_RUN_MAX:  # the growing runs: 1, 3, 7, 15 pages
            full += run
            short = np.where((short == 0) & (p <= full), full, short)
            run *= 2
        total = np.where(short > 0, short,
                         full + TRAJ_RUN_MAX * np.ceil(np.maximum(p - full, 0) / TRAJ_RUN_MAX))
        total = np.where(p <= 0, 0.0, total)
        total = total.sum() if np.ndim(total) else n * total
        return max(int(total), 1)

    @staticmethod
    def _store_fits(pages, dev):
        need = pages * (TRAJ_PAGE_STEPS * TRAJ_VALUES * 8 + 8)
        total = torch.cuda.get_device_properties(dev).total_memory
        return need <= min(TRAJ_BUDGET_BYTES, total // 3)

    def stored_steps(self, i: int) -> torch.Tensor:
        """Ray i's stored steps (accepted steps, 9) from a trajectory store:
        start state and step size, in order (for checks; walks the page
        chain on the host)."""
        if self.last_page is None:
            raise ValueError("the recording has no trajectory store")
        total = int(self.steps[i].item())
        link = self.page_link.cpu()
        pages = []
        p = int(self.last_page[i].item())
        while p >= 0:
            pages.append(p)
            p = int(link[p])
        pages.reverse()
        if len(pages) != -(-total // TRAJ_PAGE_STEPS):
            raise RuntimeError("page chain does not match the step count")
        pool = self.pool.view(-1, TRAJ_PAGE_STEPS, TRAJ_VALUES)
        idx = torch.tensor(pages, dtype=torch.int64, device=self.pool.device)
        return pool[idx].reshape(-1, TRAJ_VALUES)[:total].clone()

    @staticmethod
    def _pre_pass(st, state0, cfg):
        """The same trace without recording (statuses, steps, attempts): the
        quad tracer in the precise build, whose launch ends sooner when a
        few long rays set its time; one lane per ray in the fast build."""
        if resolve_mode(cfg.mode, cfg.rtol) == "precise":
            from .batch import TraceJob, trace_batch

            job = TraceJob(st, state0, cfg.escape_radius, cfg.max_steps, cfg.rtol, cfg.atol,
                           cfg.initial_step, cfg.max_step, cfg.min_step, cfg.capture_margin,
                           cfg.lambda_end)
            return trace_batch([job], mode="precise", tune=Tune(lanes_per_ray=4))[0]
        return trace_rays(st, state0, cfg.escape_radius, max_steps=cfg.max_steps,
                          rtol=cfg.rtol, atol=cfg.atol, initial_step=cfg.initial_step,
                          max_step=cfg.max_step, min_step=cfg.min_step,
                          capture_margin=cfg.capture_margin, mode=cfg.mode,
                          lambda_end=cfg.lambda_end)

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
        if self.last_page is not None:
            used = int(self.workspace[1].item())
            if used > self.pool_pages:
                # the estimated pool fell short: size it from a plain
                # trace's exact step counts and record again (same results)
                st, state0 = self._redo
                pre = self._pre_pass(st, state0, self.cfg)
                order = torch.argsort(pre.attempts, descending=True) if self.cfg.sort_forward \
                    else None
                used = self._pages(pre.steps, n)
                self._record_store(lib, used, order, _lib.stream_handle())
            _remember_pages(self._hint_key, used)
        elif self.cfg.trajectory is None and n:
            # checkpoints this time (the estimate did not fit): the step
            # counts size the next recording's store
            _remember_pages(self._hint_key, self._pages(self.steps, n))
        if int((self.event < 0).sum().item()) > 0:
            raise RuntimeError("a ray exceeded its checkpoint capacity; raise cap")
        gterm = torch.zeros((8, ld), dtype=torch.float64, device=dev)
        gterm[:, :n] = g_terminal.t()
        gy0 = torch.empty((8, ld), dtype=torch.float64, device=dev)
        gspin = torch.empty(ld, dtype=torch.float64, device=dev)
        # longest reversals first, rays of similar length in one wavefront
        order = torch.argsort(self.ev_step, descending=True) if self.cfg.sort_backward and n \
            else None
        if self.last_page is not None:
            _lib.check(lib.bh_trace_adjoint_trajectory_f64(
                self.params, n, self.score_radius, _lib.ptr(self.pool), self.pool_pages,
                _lib.ptr(self.page_link), _lib.ptr(self.last_page), _lib.ptr(self.steps),
                _lib.ptr(self.event), _lib.ptr(self.ev_step), _lib.ptr(self.evy1),
                _lib.ptr(self.prev), _lib.ptr(self.theta), _lib.ptr(gterm), _lib.ptr(gy0),
                _lib.ptr(gspin), _lib.ptr(order), _lib.stream_handle()), "trace_adjoint_trajectory")
            return gy0[:, :n].t().contiguous(), gspin[:n]
        per_ray = lib.bh_adjoint_scratch_doubles(self.cfg.K)
        if per_ray < 0:
            raise ValueError("checkpoint interval K out of range")
        # one block of per_ray x 64 doubles per wavefront of 64 rays
        scratch = torch.empty(per_ray * (-(-n // 64) * 64 or 64), dtype=torch.float64,
                              device=dev)
        _lib.check(lib.bh_trace_adjoint_f64(
            self.params, n, self.cfg.K, self.score_radius, _lib.ptr(self.ckpt),
            _lib.ptr(self.event), _lib.ptr(self.ev_step), _lib.ptr(self.evy1),
            _lib.ptr(self.prev), _lib.ptr(self.theta), _lib.ptr(gterm), _lib.ptr(gy0),
            _lib.ptr(gspin), _lib.ptr(scratch), _lib.ptr(order), _lib.stream_handle()),
            "trace_adjoint")
        return gy0[:, :n].t().contiguous(), gspin[:n]


# A forward that no backward will read: checkpoints only at the start (the
# terminal points and statuses are the recording's; the store is skipped).
_FORWARD_ONLY = dict(trajectory=False, K=1 << 20, cap=None)


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
    if not (torch.is_grad_enabled() and (state0.requires_grad or spin.requires_grad)):
        cfg = dataclasses.replace(cfg, **_FORWARD_ONLY)
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

def ks_radius_torch(spin, pos: torch.Tensor) -> torch.Tensor:
    """Kerr-Schild radius of positions (..., 3) (kerr.py:136-142), differentiable."""
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    a2 = spin * spin
    q = x * x + y * y + z * z - a2
    s = torch.clamp(torch.sqrt(q * q + 4.0 * a2 * z * z), min=TINY)
    return torch.sqrt(torch.clamp(0.5 * (q + s), min=TINY))


def terminal_directions(terminal: torch.Tensor, spin) -> torch.Tensor:
    """Unit vectors (sin th cos ph, sin th sin ph, cos th) of the spheroidal
    angles (th, ph) of each terminal point on its sphere r = const, from
    x + i y = (r + i a) sin th e^{i ph}, z = r cos th: the spatial part of
    the Kerr-Schild null vector l at that point (kerr.py:147-152),
    differentiable in the point and the spin."""
    pos = terminal[:, 1:4]
    r = ks_radius_torch(spin, pos)
    w = r * r + spin * spin
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    return torch.stack([(r * x + spin * y) / w, (r * y - spin * x) / w, z / r], dim=-1)


def angle_sq(n: torch.Tensor, m: torch.Tensor) -> torch.Tensor:
    """Squared great-circle angle between unit vectors, psi^2 with
    psi = 2 asin(|n - m| / 2), smooth through psi = 0 (series below
    |n - m|^2 = 1e-12)."""
    c2 = ((n - m) ** 2).sum(-1)
    big = 4.0 * torch.asin(0.5 * torch.sqrt(torch.clamp(c2, min=1e-24))) ** 2
    return torch.where(c2 < 1e-12, c2 * (1.0 + c2 / 12.0), big)


def _radial_potential(r, a, mass, lam, eta):
    """Kerr radial potential of a photon per E^4, normalised by (r^2 + a^2)^2:
    R(r) = (r^2 + a^2 - a lam)^2 - Delta (eta + (lam - a)^2),
    Delta = r^2 - 2 M r + a^2 (Carter 1968)."""
    w = r * r + a * a
    delta = r * r - 2.0 * mass * r + a * a
    return ((w - a * lam) ** 2 - delta * (eta + (lam - a) ** 2)) / (w * w)


def capture_margin(state0: torch.Tensor, spin, mass: float = 1.0, grid: int = 256,
                   refine: int = 60) -> torch.Tensor:
    """Smooth horizon-proximity score of photon states (n, 8): the minimum of
    the normalised radial potential R(r) / (r^2 + a^2)^2 over
    r_+ <= r <= r_0, from the ray's constants of motion E = -p_t,
    L_z = x p_y - y p_x and Carter's Q = p_th^2 + cos^2 th (L_z^2 / sin^2 th
    - a^2 E^2), with lam = L_z / E, eta = Q / E^2 and, in Kerr-Schild
    coordinates, p_th = cot th (x p_x + y p_y) - r sin th p_z.  An ingoing
    photon reaches the horizon iff R stays positive (score > 0): the score
    is positive inside the shadow, negative outside and zero on its edge
    (a = 0: 1 - b^2 / 27).  The minimiser r* is found on detached values
    (grid, then golden-section search); the score is R(r*) with the graph,
    whose derivative is the true one at an interior minimum (dR/dr = 0
    there), and r* = r_+(a) is kept differentiable at the horizon end."""
    a = spin if isinstance(spin, torch.Tensor) else torch.tensor(float(spin), dtype=torch.float64,
                                                                 device=state0.device)
    pos = state0[:, 1:4]
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    p_t, px, py, pz = state0[:, 4], state0[:, 5], state0[:, 6], state0[:, 7]
    r = ks_radius_torch(a, pos)
    cos_t = z / r
    sin2 = torch.clamp(1.0 - cos_t * cos_t, min=TINY)
    sin_t = torch.sqrt(sin2)
    p_th = cos_t / sin_t * (x * px + y * py) - r * sin_t * pz
    energy = -p_t
    l_z = x * py - y * px
    q = p_th * p_th + cos_t * cos_t * (l_z * l_z / sin2 - a * a * energy * energy)
    lam = l_z / energy
    eta = q / (energy * energy)
    r_plus = mass + torch.sqrt(mass * mass - a * a)
    if state0.is_cuda:
        r_star, at_horizon = _radial_minimum_device(a, mass, lam, eta, r, grid, refine)
    else:
        r_star, at_horizon = _radial_minimum_torch(a, mass, lam, eta, r_plus, r, grid, refine)
    r_eval = torch.where(at_horizon, r_plus.expand_as(r_star), r_star)
    return _radial_potential(r_eval, a, mass, lam, eta)


def _radial_minimum_device(a, mass, lam, eta, r, grid, refine):
    """capture_margin's detached search by the HIP kernel (one lane per ray;
    the spin stays on the device)."""
    lib = _lib.require_device()
    with torch.no_grad():
        lamd = lam.detach().contiguous()
        etad = eta.detach().contiguous()
        hi = r.detach().contiguous()
        ad = a.detach().to(device=lamd.device, dtype=torch.float64).reshape(1).contiguous()
        n = lamd.shape[0]
        r_star = torch.empty_like(lamd)
        at_lo = torch.empty(n, dtype=torch.int32, device=lamd.device)
        _lib.check(lib.bh_radial_minimum_f64(
            float(mass), _lib.ptr(ad), _lib.ptr(lamd), _lib.ptr(etad), _lib.ptr(hi), n, grid,
            refine, _lib.ptr(r_star), _lib.ptr(at_lo), _lib.stream_handle()), "radial_minimum")
    return r_star, at_lo.bool()


def _radial_minimum_torch(a, mass, lam, eta, r_plus, r, grid, refine):
    """capture_margin's detached search in torch operations (host tensors)."""
    with torch.no_grad():
        ad, lamd, etad = a.detach(), lam.detach(), eta.detach()
        lo_r = r_plus.detach().expand_as(r)
        # a state at or inside the horizon has no interval outside it: the
        # search then collapses onto r_+
        hi_r = torch.maximum(r.detach(), lo_r)
        u = torch.linspace(0.0, 1.0, grid, dtype=torch.float64, device=r.device)
        rr = lo_r[:, None] + (hi_r - lo_r)[:, None] * u[None, :]
        vals = _radial_potential(rr, ad, mass, lamd[:, None], etad[:, None])
        k = torch.argmin(vals, dim=1)
        lo = rr.gather(1, torch.clamp(k - 1, min=0)[:, None])[:, 0]
        hi = rr.gather(1, torch.clamp(k + 1, max=grid - 1)[:, None])[:, 0]
        g = (5.0 ** 0.5 - 1.0) / 2.0
        for _ in range(refine):
            c = hi - g * (hi - lo)
            d = lo + g * (hi - lo)
            left = (_radial_potential(c, ad, mass, lamd, etad)
                    < _radial_potential(d, ad, mass, lamd, etad))
            hi = torch.where(left, d, hi)
            lo = torch.where(left, lo, c)
        r_star = 0.5 * (lo + hi)
        at_horizon = k == 0
    return r_star, at_horizon


def edge_weights(state0: torch.Tensor, spin, scale: float, status: torch.Tensor | None = None,
                 target_status: torch.Tensor | None = None, mass: float = 1.0) -> torch.Tensor:
    """Per-ray weights (n,) for ``shadow_loss`` that fade out rays near the
    shadow's edge: min(1, |m| / scale)^2 with m the ray's ``capture_margin``,
    computed without a graph (the weights are constants of the loss).  Near
    the critical curve a ray winds around the hole before leaving, so its
    terminal direction turns through large angles for small parameter
    changes: per-ray gradients reach ~1e5 there and dominate the sum on
    fine images.  Rays whose class differs from the target's (given both
    statuses) keep weight 1: their term is the margin itself and must pull
    them across the edge."""
    if not scale > 0.0:
        raise ValueError("edge scale must be positive")
    with torch.no_grad():
        a = spin.detach() if isinstance(spin, torch.Tensor) else spin
        m = capture_margin(state0.detach(), a, mass)
        w = torch.clamp(m.abs() / scale, max=1.0) ** 2
        if status is not None and target_status is not None:
            mism = (((status == 1) & (target_status == 2))
                    | ((status == 2) & (target_status == 1)))
            w = torch.where(mism, torch.ones_like(w), w)
    return w


def shadow_loss(terminal: torch.Tensor, status: torch.Tensor, spin, target_dirs: torch.Tensor,
                target_status: torch.Tensor, state0: torch.Tensor | None = None,
                classes: tuple = (1, 2), margin_weight: float = 1.0,
                mass: float = 1.0, ray_weights: torch.Tensor | None = None) -> torch.Tensor:
    """Shadow-fit loss, summed over rays:

    * escaped in both the trace and the target: psi^2, the squared angle
      between the (th, ph) where the ray leaves through the escape sphere
      and the target's (``terminal_directions``);
    * captured in both: psi^2 between the (th, ph) where the ray plunges
      through the scoring sphere r = score_factor r_+ near the horizon and
      the target's;
    * captured in one and escaped in the other (needs ``state0``):
      ``margin_weight`` * m^2 with m the ray's ``capture_margin``, a smooth
      horizon-proximity term that vanishes on the shadow's edge and pulls
      the ray towards the target's class;
    * any other status: 0.

    ``classes`` limits the matched classes scored (1 captured, 2 escaped);
    the mismatch term is scored when both are listed and ``state0`` is
    given.  ``target_dirs`` are ``terminal_directions`` of the target's
    terminal points; optional ``ray_weights`` (n,) scale each ray's term."""
    n = terminal_directions(terminal, spin)
    zero = torch.zeros(status.shape[0], dtype=torch.float64, device=terminal.device)
    scored = torch.zeros_like(status, dtype=torch.bool)
    for c in classes:
        scored |= status == c
    same = (status == target_status) & scored
    per_ray = torch.where(same, angle_sq(n, target_dirs), zero)
    if state0 is not None and 1 in classes and 2 in classes:
        mism = (((status == 1) & (target_status == 2)) | ((status == 2) & (target_status == 1)))
        # every row is scored (no host round trip to pick the mismatched
        # ones); the others see a fixed regular state, so that no
        # non-finite margin of a row the loss ignores reaches the gradient
        regular = torch.tensor([0.0, 10.0, 0.0, 1.0, -1.0, -1.0, 0.1, 0.0],
                               dtype=state0.dtype, device=state0.device)
        m = capture_margin(torch.where(mism[:, None], state0, regular), spin, mass)
        per_ray = per_ray + torch.where(mism, margin_weight * m * m, zero)
    if ray_weights is not None:
        per_ray = per_ray * ray_weights
    return per_ray.sum()


# ---------------------------------------------------------------------------
# spin and inclination fit to a shadow image
# ---------------------------------------------------------------------------

@dataclass
class FitStep:
    """One iteration of fit_shadow: parameters and loss before the update,
    the adjoint gradients, and how many scored rays are misclassified."""

    iteration: int
    spin: float
    inclination: float
    loss: float
    grad_spin: float
    grad_inclination: float
    mismatched: int


def fit_shadow(target_dirs: torch.Tensor, target_status: torch.Tensor, spin0: float,
               inclination0: float, distance: float, fov_degrees: float, width: int,
               height: int, steps: int = 60, lr: tuple = (0.005, 0.5), gamma: float = 0.93,
               escape_radius: float = 45.0, rtol: float = 1e-8, atol: float = 1e-10,
               K: int = 32, mode: str = "auto",
               edge_scale: float | None = 0.05) -> list[FitStep]:
    """Fit spin and inclination (degrees) of a pinhole camera at ``distance``
    to a target image -- the ``terminal_directions`` and statuses of the
    target's rays, pixel for pixel -- by Adam on ``shadow_loss`` with the
    adjoint's gradients: per-parameter step sizes ``lr`` (spin,
    inclination), decayed by ``gamma`` each iteration; captures at
    r_+ + 0.01 M for the current spin, spin kept inside |a| < 0.999.  With
    ``edge_scale`` each iteration weights the rays by ``edge_weights`` at
    the current parameters (None: unweighted), which keeps the photon
    ring's windings from steering fine images.  Returns the history (the
    last entry is the final iterate's loss, weighted as it was fitted)."""
    dev = target_dirs.device
    a = torch.tensor(float(spin0), dtype=torch.float64, device=dev, requires_grad=True)
    inc = torch.tensor(float(inclination0), dtype=torch.float64, device=dev, requires_grad=True)
    opt = torch.optim.Adam([{"params": [a], "lr": lr[0]}, {"params": [inc], "lr": lr[1]}])
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=gamma)
    history: list[FitStep] = []
    for it in range(steps + 1):
        opt.zero_grad()
        spin_now = float(a.detach())
        r_plus = 1.0 + math.sqrt(1.0 - spin_now * spin_now)
        cfg = TraceSettings(escape_radius=escape_radius, rtol=rtol, atol=atol,
                            capture_margin=0.01 / r_plus, K=K, mode=mode)
        y0 = camera_state(a, inc, distance, fov_degrees, width, height, device=dev)
        Y, st = trace_terminal(y0, a, cfg)
        w = (None if edge_scale is None
             else edge_weights(y0, a, edge_scale, st, target_status))
        loss = shadow_loss(Y, st, a, target_dirs, target_status, state0=y0, ray_weights=w)
        final = it == steps
        if not final:
            loss.backward()
        mism = int((((st == 1) & (target_status == 2)) | ((st == 2) & (target_status == 1)))
                   .sum())
        history.append(FitStep(it, spin_now, float(inc.detach()), float(loss.detach()),
                               float(a.grad) if not final else float("nan"),
                               float(inc.grad) if not final else float("nan"), mism))
        if final:
            break
        opt.step()
        sched.step()
        with torch.no_grad():
            a.clamp_(-0.999, 0.999)
    return history
