"""CPU oracle: a NumPy restatement of BlackHorizon's batched adaptive tracer.

TEST INFRASTRUCTURE ONLY. This module is the checker for the HIP tracer in
``blackhorizon_amd``. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check or to
time the CPU path. The product (``blackhorizon_amd``) never imports it and has
no CPU fallback.

Every function restates one function of the reference, written from the
reference's mathematics and semantics, with the same elementwise operation
order so that IEEE rounding matches NumPy-on-NumPy (one correctly rounded
operation per ufunc call). Citations are ``file:line`` in the reference
repository (MegaKruk/BlackHorizon).

Pinning (what checks each function independently of this file):

* ``geometry``/``kerr_schild_radius``: closed forms (a = 0 gives H = 1/r and
  l = x/r; on-axis r = |z|; equatorial r = sqrt(x^2 - a^2)), l null in eta,
  g g^-1 = I, analytic gradients vs central finite differences, a = 0
  gradient closed form (tests/test_oracle_geometry.py).
* ``geodesic_rhs``: Hamilton's equations, i.e. RHS = (dH/dp, -dH/dx) by
  central differences of ``hamiltonian``; dH/dlambda = 0 along the RHS.
* DP5 tableau/step: Dormand & Prince (1980) row sums and order conditions,
  fifth-order convergence on the harmonic oscillator, embedded error band.
* ``error_ratio``/``step_factor``: closed-form known answers, NaN -> inf.
* ``trace_rays``: golden vectors written by ``tests/golden/make_golden.py``
  from the reference's own ``trace_rays`` (per-ray states, status, steps,
  iterations), the shadow-boundary closed forms (b_c = 3 sqrt 3, Bardeen
  a = 0.9 prograde/retrograde within 0.5%), and the capture mask of the
  reference-held image ``shadow_gpu.png``.
* ``null_momentum_from_velocity``/``timelike_momentum``: g^{mu nu} p p = 0 or
  -1, tangential photon at 3M has b = 3 sqrt 3.
* ``PinholeCamera``: orthonormal basis, centre ray along the target,
  horizontal FOV edge angles, row-major pixel order.
* ``integrate_dense``: conservation tolerances of the reference's orbit tests.

Float64 throughout; the only non-correctly-rounded NumPy operation is
``x ** -0.2`` in ``step_factor`` (libm/SVML ``pow``), which is reproduced by
calling the same NumPy ufunc.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

TINY = 1e-30  # kerr.py:26


# ---------------------------------------------------------------------------
# L1 geometry (kerr.py)
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class Geometry:
    """Mirror of ``KerrSchildGeometry`` (kerr.py:31-49)."""

    radius: np.ndarray
    h: np.ndarray
    l: np.ndarray
    grad_h: np.ndarray | None = None
    grad_l: np.ndarray | None = None


class Kerr:
    """Restatement of ``KerrSpacetime`` (kerr.py:52-193)."""

    def __init__(self, mass: float = 1.0, spin: float = 0.0) -> None:
        # kerr.py:68-73
        if mass <= 0.0:
            raise ValueError("mass must be positive")
        if abs(spin) > mass:
            raise ValueError("Kerr bound violated: |spin| must be <= mass")
        self.mass = float(mass)
        self.spin = float(spin)

    @property
    def outer_horizon_radius(self) -> float:  # kerr.py:75-79
        m, a = self.mass, self.spin
        return m + math.sqrt(m * m - a * a)

    @property
    def inner_horizon_radius(self) -> float:  # kerr.py:81-85
        m, a = self.mass, self.spin
        return m - math.sqrt(m * m - a * a)

    def photon_orbit_radius(self, prograde: bool = True) -> float:
        """Equatorial circular photon orbit (Bardeen 1973); kerr.py:107-112."""
        chi = abs(self.spin) / self.mass
        sign = -1.0 if prograde else 1.0
        return 2.0 * self.mass * (1.0 + math.cos((2.0 / 3.0) * math.acos(sign * chi)))

    def isco_radius(self, prograde: bool = True) -> float:
        """Bardeen-Press-Teukolsky ISCO; kerr.py:92-105."""
        chi = abs(self.spin) / self.mass
        z1 = 1.0 + (1.0 - chi * chi) ** (1.0 / 3.0) * (
            (1.0 + chi) ** (1.0 / 3.0) + (1.0 - chi) ** (1.0 / 3.0)
        )
        z2 = math.sqrt(3.0 * chi * chi + z1 * z1)
        root = math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))
        return self.mass * (3.0 + z2 + (-1.0 if prograde else 1.0) * root)

    def geometry(self, x, y, z, gradients: bool = False) -> Geometry:
        """KS radius, scalar H, null covector l and gradients (kerr.py:118-193).

        Operation order follows the reference expression by expression so
        that each intermediate is rounded exactly as NumPy rounds it there.
        """
        m, a = self.mass, self.spin
        a2 = a * a
        rho2 = x * x + y * y + z * z
        q = rho2 - a2
        s = np.sqrt(q * q + 4.0 * a2 * z * z)
        s = np.maximum(s, TINY)
        r2 = 0.5 * (q + s)
        r = np.sqrt(np.maximum(r2, TINY))
        r2_plus_a2 = r2 + a2
        qd = np.maximum(r2 * r2 + a2 * z * z, TINY)
        h = m * r * r2 / qd
        inv_r2a2 = 1.0 / np.maximum(r2_plus_a2, TINY)
        inv_r = 1.0 / r
        lx = (r * x + a * y) * inv_r2a2
        ly = (r * y - a * x) * inv_r2a2
        lz = z * inv_r
        l = np.stack([lx, ly, lz], axis=-1)
        if not gradients:
            return Geometry(radius=r, h=h, l=l)
        # implicit derivative of r (kerr.py:157-161)
        inv_s = 1.0 / s
        r_x = x * r * inv_s
        r_y = y * r * inv_s
        r_z = z * r2_plus_a2 * inv_r * inv_s
        # grad H (kerr.py:163-170)
        inv_qd2 = 1.0 / (qd * qd)
        dh_dr = m * r2 * (3.0 * a2 * z * z - r2 * r2) * inv_qd2
        dh_dz_explicit = -2.0 * m * a2 * z * r * r2 * inv_qd2
        grad_h = np.stack(
            [dh_dr * r_x, dh_dr * r_y, dh_dr * r_z + dh_dz_explicit], axis=-1
        )
        # grad l, quotient rule (kerr.py:172-189); grad_l[..., i, j] = d l_j / d x_i
        two_r = 2.0 * r
        dlx_dx = (r_x * x + r) * inv_r2a2 - lx * two_r * r_x * inv_r2a2
        dlx_dy = (r_y * x + a) * inv_r2a2 - lx * two_r * r_y * inv_r2a2
        dlx_dz = r_z * x * inv_r2a2 - lx * two_r * r_z * inv_r2a2
        dly_dx = (r_x * y - a) * inv_r2a2 - ly * two_r * r_x * inv_r2a2
        dly_dy = (r_y * y + r) * inv_r2a2 - ly * two_r * r_y * inv_r2a2
        dly_dz = r_z * y * inv_r2a2 - ly * two_r * r_z * inv_r2a2
        dlz_dx = -lz * r_x * inv_r
        dlz_dy = -lz * r_y * inv_r
        dlz_dz = (1.0 - lz * r_z) * inv_r
        row_x = np.stack([dlx_dx, dly_dx, dlz_dx], axis=-1)
        row_y = np.stack([dlx_dy, dly_dy, dlz_dy], axis=-1)
        row_z = np.stack([dlx_dz, dly_dz, dlz_dz], axis=-1)
        grad_l = np.stack([row_x, row_y, row_z], axis=-2)
        return Geometry(radius=r, h=h, l=l, grad_h=grad_h, grad_l=grad_l)

    def kerr_schild_radius(self, x, y, z):  # kerr.py:114-116
        return self.geometry(x, y, z).radius

    def metric(self, positions: np.ndarray) -> np.ndarray:
        """g_{mu nu} = eta + 2 H l l (kerr.py:201-217)."""
        geo = self.geometry(positions[..., 1], positions[..., 2], positions[..., 3])
        l4 = np.concatenate([np.ones_like(geo.h)[..., None], geo.l], axis=-1)
        eta = np.zeros(positions.shape[:-1] + (4, 4))
        eta[..., 0, 0] = -1.0
        for i in range(1, 4):
            eta[..., i, i] = 1.0
        return eta + 2.0 * geo.h[..., None, None] * (l4[..., :, None] * l4[..., None, :])

    def inverse_metric(self, positions: np.ndarray) -> np.ndarray:
        """g^{mu nu} = eta - 2 H l^mu l^nu (kerr.py:219-235)."""
        geo = self.geometry(positions[..., 1], positions[..., 2], positions[..., 3])
        l4 = np.concatenate([-np.ones_like(geo.h)[..., None], geo.l], axis=-1)
        eta = np.zeros(positions.shape[:-1] + (4, 4))
        eta[..., 0, 0] = -1.0
        for i in range(1, 4):
            eta[..., i, i] = 1.0
        return eta - 2.0 * geo.h[..., None, None] * (l4[..., :, None] * l4[..., None, :])


# ---------------------------------------------------------------------------
# L2 dynamics (geodesics.py)
# ---------------------------------------------------------------------------


def build_state(positions, momenta, t0: float = 0.0):
    """(n,8) = (t, x, y, z, p_t, p_x, p_y, p_z); geodesics.py:25-39."""
    n = positions.shape[0]
    state = np.empty((n, 8), dtype=positions.dtype)
    state[:, 0] = t0
    state[:, 1:4] = positions
    state[:, 4:8] = momenta
    return state


def geodesic_rhs(st: Kerr, state):
    """Hamilton's equations in KS coordinates; geodesics.py:42-74."""
    geo = st.geometry(state[:, 1], state[:, 2], state[:, 3], gradients=True)
    p_t = state[:, 4]
    p_s = state[:, 5:8]
    lp = -p_t + np.sum(geo.l * p_s, axis=-1)
    two_h_lp = 2.0 * geo.h * lp
    deriv = np.empty_like(state)
    deriv[:, 0] = -p_t + two_h_lp
    deriv[:, 1:4] = p_s - two_h_lp[:, None] * geo.l
    deriv[:, 4] = 0.0
    contraction = np.sum(geo.grad_l * p_s[:, None, :], axis=-1)
    deriv[:, 5:8] = geo.grad_h * (lp * lp)[:, None] + two_h_lp[:, None] * contraction
    return deriv


def hamiltonian(st: Kerr, state):
    """H_ham = 1/2 eta p p - H lp^2; geodesics.py:77-85."""
    geo = st.geometry(state[:, 1], state[:, 2], state[:, 3])
    p_t = state[:, 4]
    p_s = state[:, 5:8]
    lp = -p_t + np.sum(geo.l * p_s, axis=-1)
    eta_term = 0.5 * (-p_t * p_t + np.sum(p_s * p_s, axis=-1))
    return eta_term - geo.h * lp * lp


def conserved_quantities(state):
    """E = -p_t, L_z = x p_y - y p_x; geodesics.py:88-92."""
    return -state[:, 4], state[:, 1] * state[:, 6] - state[:, 2] * state[:, 5]


def coordinate_velocity(st: Kerr, state):
    """dx^mu/dlambda = g^{mu nu} p_nu; geodesics.py:95-106."""
    geo = st.geometry(state[:, 1], state[:, 2], state[:, 3])
    p_t = state[:, 4]
    p_s = state[:, 5:8]
    lp = -p_t + np.sum(geo.l * p_s, axis=-1)
    two_h_lp = 2.0 * geo.h * lp
    v = np.empty(state.shape[:-1] + (4,), dtype=state.dtype)
    v[:, 0] = -p_t + two_h_lp
    v[:, 1:4] = p_s - two_h_lp[:, None] * geo.l
    return v


def null_momentum_from_velocity(st: Kerr, positions, directions,
                                time_orientation: str = "past",
                                normalize: bool = True):
    """Null covector from a spatial direction; geodesics.py:109-170."""
    if time_orientation not in ("past", "future"):
        raise ValueError("time_orientation must be 'past' or 'future'")
    geo = st.geometry(positions[:, 0], positions[:, 1], positions[:, 2])
    h = geo.h
    l_dot_s = np.sum(geo.l * directions, axis=-1)
    s_dot_s = np.sum(directions * directions, axis=-1)
    g_tt = 2.0 * h - 1.0
    b_coef = 4.0 * h * l_dot_s
    c_coef = s_dot_s + 2.0 * h * l_dot_s * l_dot_s
    disc = b_coef * b_coef - 4.0 * g_tt * c_coef
    sqrt_disc = np.sqrt(np.maximum(disc, 0.0))
    root_a = (-b_coef + sqrt_disc) / (2.0 * g_tt)
    root_b = (-b_coef - sqrt_disc) / (2.0 * g_tt)
    if time_orientation == "future":
        k_t = np.maximum(root_a, root_b)
    else:
        k_t = np.minimum(root_a, root_b)
    l_dot_k = k_t + l_dot_s
    mom = np.empty(positions.shape[:-1] + (4,), dtype=positions.dtype)
    mom[:, 0] = -k_t + 2.0 * h * l_dot_k
    mom[:, 1:4] = directions + (2.0 * h * l_dot_k)[:, None] * geo.l
    if normalize:
        mom = mom / np.abs(mom[:, 0])[:, None]
    return mom


def timelike_momentum(st: Kerr, positions, coordinate_velocities):
    """Unit-mass covector from dx^i/dt; geodesics.py:173-209."""
    geo = st.geometry(positions[:, 0], positions[:, 1], positions[:, 2])
    h = geo.h
    v = coordinate_velocities
    l_dot_big_v = 1.0 + np.sum(geo.l * v, axis=-1)
    norm2 = -1.0 + np.sum(v * v, axis=-1) + 2.0 * h * l_dot_big_v * l_dot_big_v
    if bool(np.any(norm2 >= 0.0)):
        raise ValueError("coordinate velocity is not timelike at one or more positions")
    u_t = 1.0 / np.sqrt(-norm2)
    l_dot_u = u_t * l_dot_big_v
    mom = np.empty(positions.shape[:-1] + (4,), dtype=positions.dtype)
    mom[:, 0] = -u_t + 2.0 * h * l_dot_u
    mom[:, 1:4] = u_t[:, None] * v + (2.0 * h * l_dot_u)[:, None] * geo.l
    return mom


# ---------------------------------------------------------------------------
# L3 integrators (integrators.py)
# ---------------------------------------------------------------------------

# Dormand & Prince (1980) 5(4) tableau, autonomous form (integrators.py:23-54).
A21 = 1.0 / 5.0
A31, A32 = 3.0 / 40.0, 9.0 / 40.0
A41, A42, A43 = 44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0
A51, A52, A53, A54 = 19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0
A61, A62, A63, A64, A65 = (9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0,
                           49.0 / 176.0, -5103.0 / 18656.0)
B1, B3, B4, B5, B6 = 35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0
E1 = 71.0 / 57600.0
E3 = -71.0 / 16695.0
E4 = 71.0 / 1920.0
E5 = -17253.0 / 339200.0
E6 = 22.0 / 525.0
E7 = -1.0 / 40.0


def rk4_step(rhs, state, h):
    """Classical RK4 (integrators.py:57-73)."""
    hh = h[:, None]
    k1 = rhs(state)
    k2 = rhs(state + 0.5 * hh * k1)
    k3 = rhs(state + 0.5 * hh * k2)
    k4 = rhs(state + hh * k3)
    return state + (hh / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


def implicit_midpoint_step(rhs, state, h, iterations=4):
    """Implicit midpoint by fixed-point iteration (integrators.py:177-207)."""
    hh = h[:, None]
    midpoint = state + 0.5 * hh * rhs(state)
    for _ in range(iterations):
        midpoint = state + 0.5 * hh * rhs(midpoint)
    return state + hh * rhs(midpoint)


def dormand_prince_step(rhs, state, h):
    """Embedded DP5(4) step -> (fifth-order state, error); integrators.py:76-107."""
    hh = h[:, None]
    k1 = rhs(state)
    k2 = rhs(state + hh * (A21 * k1))
    k3 = rhs(state + hh * (A31 * k1 + A32 * k2))
    k4 = rhs(state + hh * (A41 * k1 + A42 * k2 + A43 * k3))
    k5 = rhs(state + hh * (A51 * k1 + A52 * k2 + A53 * k3 + A54 * k4))
    k6 = rhs(state + hh * (A61 * k1 + A62 * k2 + A63 * k3 + A64 * k4 + A65 * k5))
    inc = B1 * k1 + B3 * k3 + B4 * k4 + B5 * k5 + B6 * k6
    new_state = state + hh * inc
    k7 = rhs(new_state)
    err = hh * (E1 * k1 + E3 * k3 + E4 * k4 + E5 * k5 + E6 * k6 + E7 * k7)
    return new_state, err


def error_ratio(state, new_state, error, rtol, atol):
    """RMS of err/(atol + rtol max|y|), non-finite -> inf; integrators.py:110-117."""
    scale = atol + rtol * np.maximum(np.abs(state), np.abs(new_state))
    ratio = np.sqrt(np.mean((error / scale) ** 2, axis=-1))
    return np.where(np.isfinite(ratio), ratio, np.inf)


def step_factor(ratio):
    """clip(0.9 max(ratio, 1e-16)^-0.2, 0.2, 5); integrators.py:120-124."""
    safe = np.maximum(ratio, 1e-16)
    return np.clip(0.9 * safe ** (-0.2), 0.2, 5.0)


def step_factor_libm(ratio):
    """step_factor with the C library's pow instead of NumPy's power ufunc.

    NumPy's float64 power dispatches to a vendor SIMD pow on AVX-512 hosts
    and to libm elsewhere; the two differ by an ulp in a few percent of
    calls.  Tracing with this variant measures how far the reference's own
    results move between such hosts (its conditioning), which is the scale
    any other correct implementation can be held to on ill-conditioned rays.
    """
    import math

    safe = np.maximum(ratio, 1e-16)
    p = np.array([math.pow(float(v), -0.2) for v in np.ravel(safe)]).reshape(np.shape(safe))
    return np.clip(0.9 * p, 0.2, 5.0)


def step_factor_ulp_noise(seed: int):
    """step_factor whose pow result is moved by -1, 0 or +1 ulp at random:
    a stand-in for any faithfully rounded pow.  Tracing with it measures how
    far an ulp-level difference in the step factors moves each ray's final
    state (see step_factor_libm)."""
    rng = np.random.default_rng(seed)

    def fn(ratio):
        safe = np.maximum(ratio, 1e-16)
        p = safe ** (-0.2)
        p = np.where(rng.integers(-1, 2, size=np.shape(p)) > 0, np.nextafter(p, np.inf),
                     np.where(rng.integers(0, 2, size=np.shape(p)) > 0, np.nextafter(p, 0.0), p))
        return np.clip(0.9 * p, 0.2, 5.0)

    return fn


def integrate_dense(rhs, state0, lambda_total, rtol=1e-10, atol=1e-12,
                    initial_step=0.05, max_step=2.0, factor_fn=None):
    """Batch-shared adaptive DP5 over an affine interval; integrators.py:127-174.

    ``factor_fn(ratio) -> factor`` replaces the step-factor expression (the
    reference evaluates it with Python floats, i.e. libm pow)."""
    n = state0.shape[0]
    state = state0.copy()
    states = [state.copy()]
    elapsed = 0.0
    h_scalar = float(initial_step)
    while elapsed < lambda_total:
        h_scalar = min(h_scalar, max_step, lambda_total - elapsed)
        h = np.full((n,), h_scalar, dtype=state.dtype)
        new_state, err = dormand_prince_step(rhs, state, h)
        ratio = float(np.max(error_ratio(state, new_state, err, rtol, atol)))
        if ratio <= 1.0:
            state = new_state
            states.append(state.copy())
            elapsed += h_scalar
        if factor_fn is None:
            factor = float(min(max(0.9 * max(ratio, 1e-16) ** (-0.2), 0.2), 5.0))
        else:
            factor = float(factor_fn(np.array([ratio]))[0])
        h_scalar *= factor
        if h_scalar < 1e-14:
            raise RuntimeError("integrate_dense: step size underflow")
    return states


# ---------------------------------------------------------------------------
# L4 tracer (tracer.py)
# ---------------------------------------------------------------------------


class RayStatus(IntEnum):
    """tracer.py:21-30."""

    IN_FLIGHT = 0
    CAPTURED = 1
    ESCAPED = 2
    MAX_STEPS = 3
    FAILED = 4
    DISK = 5
    TERMINATED = 6


@dataclass(frozen=True)
class TraceResult:
    """tracer.py:33-47."""

    states: np.ndarray
    status: np.ndarray
    steps: np.ndarray
    iterations: int


def trace_rays(st: Kerr, state0, escape_radius, max_steps=20000, rtol=1e-9,
               atol=1e-12, initial_step=0.1, max_step=2.0, min_step=1e-10,
               capture_margin=1e-3, return_attempts: bool = False, factor_fn=None):
    """Active-set DP5 controller with termination; tracer.py:50-157.

    ``return_attempts`` additionally returns the per-ray attempt counts
    (accepted + rejected), which the reference does not expose but which
    equal the number of loop iterations each ray was active for.
    ``factor_fn`` replaces step_factor (see step_factor_libm).
    """
    factor_fn = step_factor if factor_fn is None else factor_fn
    n = state0.shape[0]
    capture_radius = st.outer_horizon_radius * (1.0 + capture_margin)
    states = state0.copy()
    status = np.full((n,), int(RayStatus.IN_FLIGHT), dtype=np.int32)
    steps = np.zeros((n,), dtype=np.int64)
    attempts = np.zeros((n,), dtype=np.int64)
    h = np.full((n,), float(initial_step), dtype=state0.dtype)

    def rhs(batch):
        return geodesic_rhs(st, batch)

    r0 = st.kerr_schild_radius(state0[:, 1], state0[:, 2], state0[:, 3])
    if bool(np.any(r0 >= escape_radius)):
        raise ValueError("escape_radius must exceed all initial radii")

    max_iterations = 4 * max_steps
    iterations = 0
    while iterations < max_iterations:
        active = np.nonzero(status == int(RayStatus.IN_FLIGHT))[0]
        if active.size == 0:
            break
        iterations += 1
        attempts[active] += 1
        y = states[active]
        h_a = h[active]
        y_new, err = dormand_prince_step(rhs, y, h_a)
        ratio = error_ratio(y, y_new, err, rtol, atol)
        accept = ratio <= 1.0
        y = np.where(accept[:, None], y_new, y)
        states[active] = y
        steps_a = steps[active] + accept.astype(steps.dtype)
        steps[active] = steps_a
        h_a = np.clip(h_a * factor_fn(ratio), 0.0, max_step)
        h[active] = h_a
        radius = st.kerr_schild_radius(y[:, 1], y[:, 2], y[:, 3])
        finite = np.all(np.isfinite(y), axis=-1)
        s = status[active]
        s = np.where(~finite, np.int32(RayStatus.FAILED), s)
        s = np.where(finite & (radius <= capture_radius), np.int32(RayStatus.CAPTURED), s)
        s = np.where(finite & (radius >= escape_radius), np.int32(RayStatus.ESCAPED), s)
        in_flight = s == int(RayStatus.IN_FLIGHT)
        s = np.where(in_flight & (h_a < min_step), np.int32(RayStatus.FAILED), s)
        s = np.where(in_flight & (steps_a >= max_steps), np.int32(RayStatus.MAX_STEPS), s)
        status[active] = s
    status = np.where(status == int(RayStatus.IN_FLIGHT), np.int32(RayStatus.MAX_STEPS), status)
    res = TraceResult(states=states, status=status, steps=steps, iterations=iterations)
    if return_attempts:
        return res, attempts
    return res


# ---------------------------------------------------------------------------
# L5 camera (camera.py) and image decoding (imaging.py)
# ---------------------------------------------------------------------------


def _normalize(v):
    norm = np.linalg.norm(v)
    if norm == 0.0:
        raise ValueError("zero-length vector cannot be normalized")
    return v / norm


@dataclass(frozen=True)
class PinholeCamera:
    """camera.py:24-120 (horizontal FOV, row-major pixels from top-left)."""

    position: tuple
    target: tuple = (0.0, 0.0, 0.0)
    up_hint: tuple = (0.0, 0.0, 1.0)
    fov_degrees: float = 20.0
    width: int = 320
    height: int = 240
    forward: np.ndarray = field(init=False, repr=False)
    right: np.ndarray = field(init=False, repr=False)
    up: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        if self.width <= 0 or self.height <= 0:
            raise ValueError("image dimensions must be positive")
        if not 0.0 < self.fov_degrees < 180.0:
            raise ValueError("fov_degrees must be in (0, 180)")
        fwd = _normalize(np.asarray(self.target, dtype=np.float64)
                         - np.asarray(self.position, dtype=np.float64))
        right = np.cross(fwd, np.asarray(self.up_hint, dtype=np.float64))
        if np.linalg.norm(right) < 1e-12:
            right = np.cross(fwd, np.array([0.0, 1.0, 0.0]))
        right = _normalize(right)
        up = np.cross(right, fwd)
        object.__setattr__(self, "forward", fwd)
        object.__setattr__(self, "right", right)
        object.__setattr__(self, "up", up)

    @classmethod
    def from_orbit(cls, distance, inclination_degrees, azimuth_degrees=0.0, **kw):
        if distance <= 0.0:
            raise ValueError("distance must be positive")
        inc = math.radians(inclination_degrees)
        azi = math.radians(azimuth_degrees)
        pos = (distance * math.sin(inc) * math.cos(azi),
               distance * math.sin(inc) * math.sin(azi),
               distance * math.cos(inc))
        return cls(position=pos, **kw)

    def ray_directions(self):
        half_w = math.tan(math.radians(self.fov_degrees) / 2.0)
        half_h = half_w * self.height / self.width
        cols = (np.arange(self.width) + 0.5) / self.width * 2.0 - 1.0
        rows = 1.0 - (np.arange(self.height) + 0.5) / self.height * 2.0
        u, v = np.meshgrid(cols * half_w, rows * half_h)
        d = (self.forward[None, None, :] + u[..., None] * self.right[None, None, :]
             + v[..., None] * self.up[None, None, :])
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        return d.reshape(-1, 3)

    def ray_origins(self):
        o = np.asarray(self.position, dtype=np.float64)
        return np.broadcast_to(o, (self.height * self.width, 3)).copy()


def camera_state(st: Kerr, cam: PinholeCamera):
    """render_shadow.py:66-74: past-directed null rays from the camera."""
    pos = cam.ray_origins()
    mom = null_momentum_from_velocity(st, pos, cam.ray_directions(), "past")
    return build_state(pos, mom)


def celestial_checkerboard(directions, cells=12):
    """imaging.py:27-50: light (104,128,178) where the polar and azimuthal
    cell indices have odd sum, dark (24,30,52) elsewhere."""
    z = np.clip(directions[:, 2], -1.0, 1.0)
    theta = np.arccos(z)
    phi = np.arctan2(directions[:, 1], directions[:, 0])
    i_theta = np.floor(theta / np.pi * cells).astype(np.int64)
    i_phi = np.floor((phi + np.pi) / (2.0 * np.pi) * 2 * cells).astype(np.int64)
    parity = ((i_theta + i_phi) % 2).astype(bool)
    light = np.array([104, 128, 178], dtype=np.uint8)
    dark = np.array([24, 30, 52], dtype=np.uint8)
    return np.where(parity[:, None], light, dark).astype(np.uint8)


def checker_parity(st: Kerr, states):
    """Light (True) / dark (False) cell of the escape direction; imaging.py:27-50, 77-81."""
    v = coordinate_velocity(st, states)[:, 1:4]
    norms = np.linalg.norm(v, axis=-1, keepdims=True)
    norms[norms == 0.0] = 1.0
    d = v / norms
    cells = 12
    theta = np.arccos(np.clip(d[:, 2], -1.0, 1.0))
    phi = np.arctan2(d[:, 1], d[:, 0])
    i_theta = np.floor(theta / np.pi * cells).astype(np.int64)
    i_phi = np.floor((phi + np.pi) / (2.0 * np.pi) * 2 * cells).astype(np.int64)
    return ((i_theta + i_phi) % 2).astype(bool)


def render_image(st: Kerr, states, status, width, height):
    """imaging.py:53-91: RGB uint8 (height, width, 3) of traced rays."""
    colors = {"shadow": (0, 0, 0), "dark": (24, 30, 52), "light": (104, 128, 178),
              "max_steps": (120, 32, 32), "failed": (200, 0, 200)}
    n = width * height
    pixels = np.empty((n, 3), dtype=np.uint8)
    pixels[:] = colors["failed"]
    esc = status == int(RayStatus.ESCAPED)
    pixels[status == int(RayStatus.CAPTURED)] = colors["shadow"]
    light = checker_parity(st, states[esc])
    pixels[esc] = np.where(light[:, None], colors["light"], colors["dark"]).astype(np.uint8)
    pixels[status == int(RayStatus.MAX_STEPS)] = colors["max_steps"]
    return pixels.reshape(height, width, 3)

