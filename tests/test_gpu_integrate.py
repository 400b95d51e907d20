"""GPU: integration to a fixed affine parameter (timelike orbits, proper
time) against the reference's integrate_dense run per ray (golden vectors),
batched jobs against plain launches, and the adjoint with an affine end
against the autograd oracle."""

import math

import numpy as np
import pytest
import torch

from conftest import golden
from oracle import ad_oracle as ado
from oracle import bh_oracle as orc
from parity import assert_trace_parity, block_diff
from blackhorizon_amd import workloads

pytestmark = pytest.mark.gpu


def _golden_cases():
    g = golden("dense_orbits.npz")
    k = 0
    out = []
    while f"c{k}_state0" in g.files:
        out.append((g[f"c{k}_state0"], float(g[f"c{k}_lambda"]), float(g[f"c{k}_rtol"]),
                    g[f"c{k}_final"], int(g[f"c{k}_steps"])))
        k += 1
    return out


def _oracle_spread(s0, lam, rtol, final):
    """How far integrate_dense's own final state moves under ulp-level step
    factor changes (NumPy SIMD pow, random-ulp pow) -- its conditioning."""
    st = orc.Kerr(spin=0.9)
    worst = 0.0
    for fn in (orc.step_factor, orc.step_factor_ulp_noise(1), orc.step_factor_ulp_noise(2)):
        alt = orc.integrate_dense(lambda y: orc.geodesic_rhs(st, y), s0, lam, rtol=rtol,
                                  factor_fn=fn)[-1]
        worst = max(worst, float(block_diff(alt, final)[0]))
    return worst


@pytest.mark.parametrize("mode", ["auto", "precise"])
def test_integrate_rays_matches_reference_integrate_dense(hip, mode):
    from blackhorizon_amd.tracer import integrate_rays

    st = hip.KerrSpacetime(spin=0.9)
    cases = _golden_cases()
    assert len(cases) >= 8
    for s0, lam, rtol, final, steps in cases:
        res = integrate_rays(st, s0, lam, rtol=rtol, mode=mode)
        assert int(res.status[0]) == 6, res.status  # TERMINATED at lambda_total
        spread = _oracle_spread(s0, lam, rtol, final)
        assert_trace_parity(res.states, res.status, res.steps, final, np.array([6]),
                            np.array([steps]), label=f"lambda {lam}", spread=np.array([spread]))


def test_integrate_rays_conservation(hip):
    """test_orbits.py tolerances on the final states: E exact, L_z 1e-8
    (photon flyby, rtol 1e-11) and H within 1e-9; the r = 6 M circular orbit
    stays circular and in the equatorial plane."""
    from blackhorizon_amd.tracer import integrate_rays

    st = hip.KerrSpacetime(spin=0.9)
    ost = orc.Kerr(spin=0.9)
    cases = _golden_cases()
    s0, lam, rtol, _, _ = cases[0]
    res = integrate_rays(st, s0, lam, rtol=rtol)
    e0, lz0 = orc.conserved_quantities(s0)
    e1, lz1 = orc.conserved_quantities(res.states)
    np.testing.assert_array_equal(e1, e0)
    np.testing.assert_allclose(lz1, lz0, rtol=1e-8)
    np.testing.assert_allclose(orc.hamiltonian(ost, res.states), 0.0, atol=1e-9)
    s0, lam, rtol, _, _ = cases[1]
    res = integrate_rays(st, s0, lam, rtol=rtol)
    y = res.states
    np.testing.assert_allclose(ost.kerr_schild_radius(y[:, 1], y[:, 2], y[:, 3]), 6.0, rtol=1e-6)
    assert abs(y[0, 3]) < 1e-8
    np.testing.assert_allclose(orc.hamiltonian(ost, y), -0.5, rtol=1e-9)


def _orbit_states(job):
    st = orc.Kerr(spin=job.spin)
    return orc.build_state(job.positions,
                           orc.timelike_momentum(st, job.positions, job.velocities))


def test_batched_orbit_jobs_equal_plain(hip):
    from blackhorizon_amd import batch
    from blackhorizon_amd.tracer import integrate_rays

    jobs = workloads.orbit_batches(n_total=600, n_jobs=5, lambda_end=150.0, rtol=1e-10)
    tj = [batch.TraceJob(hip.KerrSpacetime(spin=j.spin), _orbit_states(j), j.escape_radius,
                         max_steps=j.max_steps, rtol=j.rtol, atol=j.atol,
                         initial_step=j.initial_step, max_step=j.max_step, min_step=j.min_step,
                         capture_margin=j.capture_margin, lambda_end=j.lambda_end)
          for j in jobs]
    res = batch.trace_batch(tj)
    for j, t, r in zip(jobs, tj, res):
        alone = integrate_rays(t.spacetime, t.state0, j.lambda_end, rtol=j.rtol, atol=j.atol,
                               initial_step=j.initial_step, max_step=j.max_step,
                               min_step=j.min_step, max_steps=j.max_steps,
                               escape_radius=j.escape_radius, capture_margin=j.capture_margin)
        np.testing.assert_array_equal(r.status, alone.status)
        np.testing.assert_array_equal(r.steps, alone.steps)
        np.testing.assert_array_equal(r.states, alone.states)
        # perturbed orbits that start near the ISCO may plunge
        assert np.all((r.status == 6) | (r.status == 1))
        assert np.mean(r.status == 6) > 0.9


def test_adjoint_with_affine_end_matches_autograd_oracle(hip):
    """Gradients of final positions after proper time 60 M with respect to
    the initial momenta and the spin (timelike orbits, no terminal event)."""
    from blackhorizon_amd import adjoint

    p, v = workloads.near_circular_orbits(8, seed=21)
    s0 = _orbit_states(workloads.OrbitJob(spin=0.9, positions=p, velocities=v,
                                          lambda_end=60.0, rtol=1e-10))
    kw = dict(escape_radius=1000.0, rtol=1e-10, atol=1e-12, initial_step=0.05, min_step=1e-14,
              lambda_end=60.0)
    w = np.random.default_rng(6).normal(size=(8, 8))
    w[:, [0, 4]] = 0.0
    so = torch.from_numpy(s0).requires_grad_(True)
    ao = torch.tensor(0.9, dtype=torch.float64, requires_grad=True)
    Yo, sto, _ = ado.trace_terminal(so, ao, max_steps=10_000_000, **kw)
    assert np.all(sto.numpy() == 6)
    go_s, go_a = torch.autograd.grad((Yo * torch.from_numpy(w)).sum(), [so, ao])
    cfg = adjoint.TraceSettings(max_steps=10_000_000, K=16, **kw)
    sg = torch.from_numpy(s0).cuda().requires_grad_(True)
    ag = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
    Yg, stg = adjoint.trace_terminal(sg, ag, cfg)
    assert np.all(stg.cpu().numpy() == 6)
    (Yg * torch.from_numpy(w).cuda()).sum().backward()
    ref = go_s.numpy()
    scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-12)
    assert np.max(np.abs(sg.grad.cpu().numpy() - ref) / scale) < 1e-6
    assert abs(float(ag.grad) - float(go_a)) <= 1e-6 * abs(float(go_a))


# ---------------------------------------------------------------------------
# integrate_dense: the batch-shared step sequence (integrators.py:127-174)
# ---------------------------------------------------------------------------

def _dense_batch(n=6, seed=11):
    st = orc.Kerr(spin=0.9)
    p, v = workloads.near_circular_orbits(n, seed=seed)
    return orc.build_state(p, orc.timelike_momentum(st, p, v))


def test_integrate_dense_single_ray_matches_golden(hip):
    """n = 1: the batch maximum is the ray's own ratio, so the golden
    integrate_dense runs pin the whole accepted-step count and final state."""
    from blackhorizon_amd.integrators import integrate_dense

    st = hip.KerrSpacetime(spin=0.9)
    for k, (s0, lam, rtol, final, steps) in enumerate(_golden_cases()):
        out = integrate_dense(st, s0, lam, rtol=rtol)
        assert len(out) == steps + 1, (k, len(out), steps)
        np.testing.assert_array_equal(out[0], s0)
        spread = _oracle_spread(s0, lam, rtol, final)
        d = float(block_diff(out[-1], final)[0])
        assert d <= max(1e-7, 10.0 * spread), (k, d, spread)


@pytest.mark.parametrize("mode,rtol,tol", [("precise", 1e-10, 1e-12), ("fast", 1e-8, 1e-6)])
def test_integrate_dense_batch_matches_oracle(hip, mode, rtol, tol):
    """Six orbits sharing one step sequence: same number of accepted states
    and every state of the trajectory close to the oracle's.  precise: the
    reference's rounding except the pow of the step factor, so 1e-12, or 10x
    the spread an ulp-level change of the step factors (NumPy's pow,
    random-ulp pow) causes in the oracle's own trajectory.  fast (opt-in
    here; integrate_dense defaults to precise): FMA rounding of the stages
    moves the embedded error estimate, whose size is ~ratio * rtol, by
    ~1e-16 relative to the state, hence the shared step sizes by up to
    ~1e-16 / (ratio * rtol); measured 1.2e-7 on the coordinate-time block
    (positions and momenta ~1e-8), bar 1e-6."""
    from blackhorizon_amd.integrators import GeodesicRHS, integrate_dense

    s0 = _dense_batch()
    ost = orc.Kerr(spin=0.9)

    def oracle(fn=None):
        return orc.integrate_dense(lambda y: orc.geodesic_rhs(ost, y), s0, 40.0, rtol=rtol,
                                   factor_fn=fn)

    ref = oracle()
    spread = np.zeros(len(ref))
    for fn in (orc.step_factor, orc.step_factor_ulp_noise(1), orc.step_factor_ulp_noise(2)):
        alt = oracle(fn)
        if len(alt) == len(ref):
            spread = np.maximum(spread, [float(np.max(block_diff(a, b)))
                                         for a, b in zip(alt, ref)])
    out = integrate_dense(GeodesicRHS(hip.KerrSpacetime(spin=0.9), mode), s0, 40.0, rtol=rtol)
    assert len(out) == len(ref)
    for k, (a, b) in enumerate(zip(out, ref)):
        assert float(np.max(block_diff(a, b))) <= max(tol, 10.0 * spread[k]), k


def test_integrate_dense_chunks_resume_exactly(hip):
    from blackhorizon_amd.integrators import integrate_dense

    s0 = _dense_batch(n=3, seed=12)
    st = hip.KerrSpacetime(spin=0.9)
    whole = integrate_dense(st, s0, 25.0)
    parts = integrate_dense(st, s0, 25.0, chunk=7)
    assert len(whole) == len(parts) > 8
    for a, b in zip(whole, parts):
        np.testing.assert_array_equal(a, b)


def test_integrate_dense_edges(hip):
    from blackhorizon_amd.integrators import integrate_dense

    st = hip.KerrSpacetime(spin=0.9)
    s0 = _dense_batch(n=2)
    out = integrate_dense(st, s0, 0.0)
    assert len(out) == 1 and np.array_equal(out[0], s0)
    with pytest.raises(ValueError):
        integrate_dense(st, s0[:0], 1.0)
    bad = s0.copy()
    bad[1, 5] = np.nan  # every ratio is +inf: the step shrinks to underflow, as in the reference
    with pytest.raises(RuntimeError, match="underflow"):
        integrate_dense(st, bad, 10.0)
    # tensors in, tensors out (on the device)
    out = integrate_dense(st, torch.from_numpy(s0).cuda(), 5.0)
    assert isinstance(out[-1], torch.Tensor) and out[-1].is_cuda and out[-1].shape == (2, 8)


# ---------------------------------------------------------------------------
# rk4_step (integrators.py:57-73), the reference benchmark's stepper
# ---------------------------------------------------------------------------

def _rk4_inputs():
    st = orc.Kerr(spin=0.9)
    rng = np.random.default_rng(7)
    u = rng.normal(size=(300, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    photons = orc.build_state(50.0 * u, orc.null_momentum_from_velocity(st, 50.0 * u, -u))
    orbits = _dense_batch(n=60, seed=13)
    s = np.concatenate([photons, orbits])
    h = np.exp(rng.uniform(np.log(1e-3), np.log(0.5), size=s.shape[0]))
    return st, s, h


def test_rk4_step_precise_bitwise_equals_oracle(hip):
    from blackhorizon_amd.integrators import GeodesicRHS, rk4_step, rk4_steps

    ost, s, h = _rk4_inputs()
    rhs = GeodesicRHS(hip.KerrSpacetime(spin=0.9), "precise")
    ref = orc.rk4_step(lambda y: orc.geodesic_rhs(ost, y), s, h)
    np.testing.assert_array_equal(rk4_step(rhs, s, h), ref)
    # n fused steps == n applications
    ref = s
    for _ in range(20):
        ref = orc.rk4_step(lambda y: orc.geodesic_rhs(ost, y), ref, h)
    np.testing.assert_array_equal(rk4_steps(rhs, s, h, 20), ref)
    np.testing.assert_array_equal(rk4_steps(rhs, s, h, 0), s)


def test_rk4_step_fast_close_to_oracle(hip):
    from blackhorizon_amd.integrators import GeodesicRHS, rk4_step

    ost, s, h = _rk4_inputs()
    got = rk4_step(GeodesicRHS(hip.KerrSpacetime(spin=0.9), "fast"), s, h)
    ref = orc.rk4_step(lambda y: orc.geodesic_rhs(ost, y), s, h)
    assert float(np.max(block_diff(got, ref))) < 1e-13


# ---------------------------------------------------------------------------
# implicit midpoint (integrators.py:177-207) and tidal-disruption debris
# ---------------------------------------------------------------------------

def _orbit_state_14():
    """test_offline.py:37-44: a = 0.9, r = 14 M, dx/dt = (0, 0.21, 0.02)."""
    st = orc.Kerr(spin=0.9)
    pos = np.array([[14.0, 0.0, 0.0]])
    return st, orc.build_state(pos, orc.timelike_momentum(st, pos, np.array([[0.0, 0.21, 0.02]])))


def test_implicit_midpoint_bitwise_equals_oracle(hip):
    from blackhorizon_amd.integrators import GeodesicRHS, implicit_midpoint_step, \
        implicit_midpoint_steps

    ost, s, h = _rk4_inputs()
    rhs = GeodesicRHS(hip.KerrSpacetime(spin=0.9), "precise")
    f = lambda y: orc.geodesic_rhs(ost, y)  # noqa: E731
    np.testing.assert_array_equal(implicit_midpoint_step(rhs, s, h),
                                  orc.implicit_midpoint_step(f, s, h))
    ref = s
    for _ in range(10):
        ref = orc.implicit_midpoint_step(f, ref, h, iterations=6)
    np.testing.assert_array_equal(implicit_midpoint_steps(rhs, s, h, 10, iterations=6), ref)


def test_implicit_midpoint_second_order(hip):
    """test_offline.py:46-67: halving the step cuts the one-orbit position
    error about 4x."""
    from blackhorizon_amd.integrators import implicit_midpoint_steps

    _, s0 = _orbit_state_14()
    st = hip.KerrSpacetime(spin=0.9)

    def endpoint(h, steps):
        return implicit_midpoint_steps(st, s0, np.array([h]), steps, iterations=6)

    ref = endpoint(0.025, 4000)
    e_coarse = np.abs(endpoint(0.2, 500)[0, 1:4] - ref[0, 1:4]).max()
    e_fine = np.abs(endpoint(0.1, 1000)[0, 1:4] - ref[0, 1:4]).max()
    assert 1.6 < np.log2(e_coarse / e_fine) < 2.6


def test_midpoint_energy_bounded_where_rk4_drifts(hip):
    """test_offline.py:69-107: 30,000 steps of h = 0.6; the RK4 Hamiltonian
    error grows secularly, the midpoint one oscillates and stays bounded."""
    from blackhorizon_amd.integrators import implicit_midpoint_steps, rk4_steps

    ost, s0 = _orbit_state_14()
    st = hip.KerrSpacetime(spin=0.9)
    h = np.array([0.6])
    samples = {}
    for name, stepper in (("rk4", rk4_steps), ("midpoint", implicit_midpoint_steps)):
        state, errors = s0, []
        for _ in range(12):
            state = stepper(st, state, h, 2500)
            errors.append(abs(float(orc.hamiltonian(ost, state)[0]) + 0.5))
        samples[name] = np.asarray(errors)
    rk4 = samples["rk4"]
    assert np.all(np.diff(rk4) > 0.0) and rk4[-1] > 5.0 * rk4[0]
    mid = samples["midpoint"]
    assert float(mid[-1]) < 0.05 * float(mid.max()) and float(mid.max()) < 1e-2


def test_debris_stream_energy_spread(hip):
    """test_dynamics.py:211-221."""
    from blackhorizon_amd import dynamics

    st = hip.KerrSpacetime(spin=0.9)
    stream = dynamics.generate_debris_stream(st, star_mass=1e-6, star_radius=0.5,
                                             n_particles=4000, seed=3)
    assert 0.3 < stream.bound_fraction < 0.7
    spread = float(np.std(stream.specific_energies))
    analytic = dynamics.energy_spread_estimate(1.0, 0.5, stream.tidal_radius)
    assert 0.1 * analytic < spread < 3.0 * analytic
    # the centre moves at the marginally bound speed: E = 1 there
    v = dynamics.marginally_bound_speed(st, [stream.tidal_radius, 0.0, 0.0], [0.0, 1.0, 0.0])
    ost = orc.Kerr(spin=0.9)
    p = orc.timelike_momentum(ost, np.array([[stream.tidal_radius, 0.0, 0.0]]),
                              np.array([[0.0, v, 0.0]]))
    assert abs(-p[0, 0] - 1.0) < 1e-12


def test_debris_evolves_on_valid_geodesics(hip):
    """test_dynamics.py:223-236, 400 RK4 steps of h = 0.5 in one launch."""
    from blackhorizon_amd import dynamics
    from blackhorizon_amd.integrators import rk4_steps

    st = hip.KerrSpacetime(spin=0.9)
    stream = dynamics.generate_debris_stream(st, star_mass=1e-6, star_radius=0.5,
                                             n_particles=64, seed=5, device_states=True)
    out = rk4_steps(st, stream.states, torch.full((64,), 0.5, dtype=torch.float64,
                                                  device="cuda"), 400)
    np.testing.assert_allclose(orc.hamiltonian(orc.Kerr(spin=0.9), out.cpu().numpy()), -0.5,
                               atol=1e-5)


# ---------------------------------------------------------------------------
# the reference's orbit checks (test_orbits.py) on device trajectories
# ---------------------------------------------------------------------------

def _trajectory(hip, spin, s0, lam, rtol):
    from blackhorizon_amd import integrators

    states = integrators.integrate_dense(hip.KerrSpacetime(spin=spin), s0, lambda_total=lam,
                                         rtol=rtol)
    return np.concatenate(states, axis=0)


def test_photon_sphere_orbit_on_device(hip):
    """test_orbits.py:68-91: a tangential photon at r = 3 M (a = 0) rides
    the unstable circular orbit for 1.5 turns with r within 1e-6 of 3 M at
    every accepted step of integrate_dense at rtol 1e-12, and winds by more
    than 2.5 pi; its impact parameter is 3 sqrt 3 (test_orbits.py:93-104)."""
    st = orc.Kerr(spin=0.0)
    pos = np.array([[3.0, 0.0, 0.0]])
    mom = orc.null_momentum_from_velocity(st, pos, np.array([[0.0, 1.0, 0.0]]), "future")
    s0 = orc.build_state(pos, mom)
    e, lz = orc.conserved_quantities(s0)
    np.testing.assert_allclose(lz[0] / e[0], 3.0 * math.sqrt(3.0), rtol=1e-12)
    traj = _trajectory(hip, 0.0, s0, 16.0, 1e-12)
    r = st.kerr_schild_radius(traj[:, 1], traj[:, 2], traj[:, 3])
    print(f"photon sphere: {traj.shape[0] - 1} steps, max |r - 3| {np.max(np.abs(r - 3.0)):.2e}")
    np.testing.assert_allclose(r, 3.0, atol=1e-6)
    phi = np.unwrap(np.arctan2(traj[:, 2], traj[:, 1]))
    assert abs(phi[-1] - phi[0]) > 2.5 * np.pi


def test_reference_orbit_checks_on_device_trajectories(hip):
    """test_orbits.py:22-64 and 108-136 on every accepted state of the
    device's integrate_dense: the a = 0.9 strong-field photon flyby (E to
    1e-12, L_z to 1e-8, H to 1e-9, closest approach below 6 M), the
    timelike normalisation H = -1/2 to 1e-9, and the r = 6 M prograde
    circular orbit (r to 1e-6, z to 1e-8, E to 1e-12, L_z to 1e-10)."""
    st = orc.Kerr(spin=0.9)
    pos = np.array([[30.0, 4.0, 1.5]])
    mom = orc.null_momentum_from_velocity(st, pos, np.array([[-1.0, 0.0, 0.0]]), "future")
    traj = _trajectory(hip, 0.9, orc.build_state(pos, mom), 70.0, 1e-11)
    e, lz = orc.conserved_quantities(traj)
    np.testing.assert_allclose(e, e[0], rtol=1e-12)
    np.testing.assert_allclose(lz, lz[0], rtol=1e-8)
    np.testing.assert_allclose(orc.hamiltonian(st, traj), 0.0, atol=1e-9)
    assert float(st.kerr_schild_radius(traj[:, 1], traj[:, 2], traj[:, 3]).min()) < 6.0
    om = 1.0 / (6.0 ** 1.5 + 0.9)
    pos = np.array([[6.0, 0.9, 0.0]])
    s0 = orc.build_state(pos, orc.timelike_momentum(st, pos, np.array([[-om * 0.9, om * 6.0,
                                                                          0.0]])))
    traj = _trajectory(hip, 0.9, s0, 120.0, 1e-11)
    np.testing.assert_allclose(orc.hamiltonian(st, traj), -0.5, rtol=1e-9)
    traj = _trajectory(hip, 0.9, s0, 170.0, 1e-11)
    np.testing.assert_allclose(st.kerr_schild_radius(traj[:, 1], traj[:, 2], traj[:, 3]), 6.0,
                               rtol=1e-6)
    assert np.allclose(traj[:, 3], 0.0, atol=1e-8)
    e, lz = orc.conserved_quantities(traj)
    np.testing.assert_allclose(e, e[0], rtol=1e-12)
    np.testing.assert_allclose(lz, lz[0], rtol=1e-10)


# ---------------------------------------------------------------------------
# the timelike-orbit workload at its configured size
# ---------------------------------------------------------------------------

def test_timelike_orbit_workload_full_size(hip):
    """65,536 near-circular orbits around a = 0.9 in 16 uneven jobs,
    integrated to proper time 1e4 M at rtol 1e-10 in one persistent batched
    plan (workloads.orbit_batches() defaults):

    * every job's results equal one plain launch of that job, bit for bit;
    * p_t is carried exactly (E drift 0);
    * every orbit ends TERMINATED at tau = 1e4, CAPTURED, or ESCAPED, and
      only orbits with E > 0.999 (unbound or apoapsis beyond ~1000 M)
      escape; at least 97% reach tau;
    * on the orbits that reach tau: relative L_z drift median below 1e-10
      and max below 1e-5, |H + 1/2| drift median below 1e-11 and max below
      1e-6 (measured: 2.9e-12 / 7.8e-7 and 1.7e-13 / 2.6e-8);
    * seeded orbits inside the launch meet the parity bar against the
      reference's integrate_dense run on each alone (c5_orbits.npz), within
      10x the reference's own host-to-host spread where that exceeds 1e-7.
    """
    import time

    import torch

    from blackhorizon_amd import batch, geodesics, kerr
    from blackhorizon_amd.tracer import integrate_rays

    jobs = workloads.orbit_batches()
    assert len(jobs) == 16 and sum(j.n_rays for j in jobs) == 65536
    st = kerr.KerrSpacetime(1.0, 0.9)
    tjobs, states = [], []
    for j in jobs:
        pos = torch.from_numpy(j.positions).cuda()
        mom = geodesics.timelike_momentum(st, pos, torch.from_numpy(j.velocities).cuda())
        s0 = geodesics.build_state(pos, mom)
        states.append(s0)
        tjobs.append(batch.TraceJob(st, s0, j.escape_radius, j.max_steps, j.rtol, j.atol,
                                    j.initial_step, j.max_step, j.min_step, j.capture_margin,
                                    j.lambda_end))
    plan = batch.BatchPlan(tjobs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.launch()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    res = plan.results(as_numpy=False)
    s0 = torch.cat(states)
    s1 = torch.cat([r.states for r in res])
    status = torch.cat([r.status for r in res]).cpu().numpy()
    t0 = time.perf_counter()
    for j, t, r in zip(jobs, tjobs, res):
        alone = integrate_rays(st, t.state0, j.lambda_end, rtol=j.rtol, atol=j.atol,
                               initial_step=j.initial_step, max_step=j.max_step,
                               min_step=j.min_step, max_steps=j.max_steps,
                               escape_radius=j.escape_radius, capture_margin=j.capture_margin)
        assert torch.equal(r.status, alone.status)
        assert torch.equal(r.steps, alone.steps)
        assert torch.equal(r.states, alone.states)
    t_plain = time.perf_counter() - t0
    assert torch.equal(s1[:, 4], s0[:, 4])
    counts = np.bincount(status, minlength=7)
    assert counts[1] + counts[2] + counts[6] == 65536, counts
    assert counts[6] >= 0.97 * 65536
    energy = -s0[:, 4].cpu().numpy()
    assert np.all(energy[status == 2] > 0.999)
    d0 = geodesics._diagnostics(st, s0, ("ham", "lz"))
    d1 = geodesics._diagnostics(st, s1, ("ham", "lz"))
    ok = torch.from_numpy(status == 6).cuda()
    dlz = ((d1["lz"] - d0["lz"]).abs() / d0["lz"].abs())[ok]
    dh = (d1["ham"] - d0["ham"]).abs()[ok]
    steps = int(sum(int(r.steps.sum().item()) for r in res))
    print(f"orbits: {steps} steps, batched {t_batched * 1e3:.0f} ms (wall, one launch), 16 plain "
          f"launches {t_plain * 1e3:.0f} ms; statuses {counts.tolist()}; L_z drift median "
          f"{float(dlz.median()):.2e} max {float(dlz.max()):.2e}; H drift median "
          f"{float(dh.median()):.2e} max {float(dh.max()):.2e}")
    assert float(dlz.median()) < 1e-10 and float(dlz.max()) < 1e-5
    assert float(dh.median()) < 1e-11 and float(dh.max()) < 1e-6
    g = golden("c5_orbits.npz")
    k = 0
    all_final = s1.cpu().numpy()
    all_steps = torch.cat([r.steps for r in res]).cpu().numpy()
    all_s0 = s0.cpu().numpy()
    while f"o{k}_index" in g.files:
        i = int(g[f"o{k}_index"])
        np.testing.assert_array_equal(all_s0[i:i + 1], g[f"o{k}_state0"])
        assert_trace_parity(all_final[i:i + 1], status[i:i + 1], all_steps[i:i + 1],
                            g[f"o{k}_final"], np.array([6]), np.array([int(g[f"o{k}_steps"])]),
                            label=f"orbit {i}", spread=np.array([float(g[f"o{k}_spread"])]),
                            spread_status=(6,))
        k += 1
    assert k >= 6


def test_timelike_orbit_workload_adjoint_full_size(hip):
    """Gradients of the final positions of all 65,536 workload orbits with
    respect to the spin and every initial state (K = 32): finite for every
    orbit that reaches tau = 1e4; per orbit bitwise equal to a launch of a
    few of those orbits alone (trajectory store) and to the same orbits
    recorded with checkpoints; and for sampled orbits equal to central
    differences (h = 1e-6 in p_x, p_y) of the final position to 1e-5."""
    import torch

    from blackhorizon_amd import adjoint, geodesics, kerr

    jobs = workloads.orbit_batches()
    st = kerr.KerrSpacetime(1.0, 0.9)
    pos = torch.from_numpy(np.concatenate([j.positions for j in jobs])).cuda()
    vel = torch.from_numpy(np.concatenate([j.velocities for j in jobs])).cuda()
    s0 = geodesics.build_state(pos, geodesics.timelike_momentum(st, pos, vel))
    cfg = adjoint.TraceSettings(escape_radius=1000.0, max_steps=10_000_000, rtol=1e-10,
                                atol=1e-12, initial_step=0.05, min_step=1e-14, K=32,
                                lambda_end=1.0e4)
    w = torch.randn((s0.shape[0], 3), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(3))
    sg = s0.clone().requires_grad_(True)
    a = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
    Y, status = adjoint.trace_terminal(sg, a, cfg)
    (Y[:, 1:4] * w).sum().backward()
    ok = status == 6
    assert int(ok.sum()) >= 0.97 * s0.shape[0]
    assert bool(torch.isfinite(sg.grad[ok]).all()) and math.isfinite(float(a.grad))
    pick = torch.tensor([22919, 14205, 8749, 46395, 100, 40000], device="cuda")
    assert bool(ok[pick].all())
    sp = s0[pick].clone().requires_grad_(True)
    Yp, _ = adjoint.trace_terminal(sp, 0.9, cfg)
    (Yp[:, 1:4] * w[pick]).sum().backward()
    assert torch.equal(sp.grad, sg.grad[pick])
    # the whole workload is recorded into a trajectory store; checkpoints
    # and their replay give the same bits
    import dataclasses

    sc = s0[pick].clone().requires_grad_(True)
    Yc, _ = adjoint.trace_terminal(sc, 0.9, dataclasses.replace(cfg, trajectory=False))
    (Yc[:, 1:4] * w[pick]).sum().backward()
    assert torch.equal(sc.grad.view(torch.int64), sg.grad[pick].view(torch.int64))
    h = 1e-6
    for i in range(3):
        for j in (5, 6):
            d = torch.zeros_like(s0[pick[i:i + 1]])
            d[0, j] = h
            with torch.no_grad():
                rp = adjoint.Recorded(s0[pick[i:i + 1]] + d, 0.9, cfg)
                rm = adjoint.Recorded(s0[pick[i:i + 1]] - d, 0.9, cfg)
            assert int(rp.steps[0]) == int(rm.steps[0])
            fd = float(((rp.terminal()[0, 1:4] - rm.terminal()[0, 1:4]) * w[pick[i]]).sum()) / (
                2 * h)
            g = float(sp.grad[i, j])
            print(f"orbit {int(pick[i])} d/dp_{j - 4}: fd {fd:.10g} adjoint {g:.10g}")
            assert abs(fd - g) <= 1e-5 * abs(fd) + 1e-8, (int(pick[i]), j, fd, g)


@pytest.mark.parametrize("mode", ["precise", "fast"])
@pytest.mark.parametrize("n,lam", [(3000, 60.0), (140000, 6.0)])
def test_integrate_dense_independent_of_ray_placement(hip, mode, n, lam):
    """The batch-shared step comes from the largest ratio over the whole
    batch, whichever block and thread each ray lands in: orbits integrated
    in a shuffled order give the same trajectories bit for bit, and so does
    the batch split into chunks of 7 accepted states resumed from the saved
    controller state.  3,000 orbits keep one ray per thread in registers;
    140,000 exceed the chip's resident threads (131,072), so threads walk
    several rays held in the SoA work planes."""
    from blackhorizon_amd import integrators

    p, v = workloads.near_circular_orbits(n, seed=3, r_min=15.0, spread=0.02)
    st = orc.Kerr(spin=0.9)
    s0 = orc.build_state(p, orc.timelike_momentum(st, p, v))
    g = integrators.GeodesicRHS(hip.KerrSpacetime(spin=0.9), mode=mode)
    a = np.stack(integrators.integrate_dense(g, s0, lam, rtol=1e-10))
    perm = np.random.default_rng(8).permutation(n)
    b = np.stack(integrators.integrate_dense(g, s0[perm], lam, rtol=1e-10))
    assert a.shape[0] > 5
    np.testing.assert_array_equal(b[:, np.argsort(perm)], a)
    c = np.stack(integrators.integrate_dense(g, s0, lam, rtol=1e-10, chunk=7))
    np.testing.assert_array_equal(c, a)
    if n > 100000 and mode == "precise":
        # the oracle's batch integrator on the same 140,000 rows (its power
        # may differ from the device's correctly rounded one by an ulp)
        ref = orc.integrate_dense(lambda y: orc.geodesic_rhs(st, y), s0, lam, rtol=1e-10)
        assert len(ref) == a.shape[0]
        worst = float(block_diff(a[-1], ref[-1]).max())
        assert worst <= 1e-12, worst


def test_timelike_orbit_workload_fast_build_within_the_bar(hip):
    """The fast build on the whole 65,536-orbit workload (rtol 1e-10, where
    "auto" picks the precise build): statuses identical to the precise
    build's, every final state within the 1e-7 parity bar of the precise
    result (measured max 8.8e-8), and the golden orbits within 1e-7 of the
    reference's integrate_dense (measured <= 2e-11).  Bound orbits do not
    show the ~1e-16 / rtol drift that near-critical photons do."""
    import torch

    from blackhorizon_amd import batch, geodesics, kerr
    from parity import STATE_RTOL

    jobs = workloads.orbit_batches()
    st = kerr.KerrSpacetime(1.0, 0.9)
    tjobs = []
    for j in jobs:
        pos = torch.from_numpy(j.positions).cuda()
        s0 = geodesics.build_state(pos, geodesics.timelike_momentum(
            st, pos, torch.from_numpy(j.velocities).cuda()))
        tjobs.append(batch.TraceJob(st, s0, j.escape_radius, j.max_steps, j.rtol, j.atol,
                                    j.initial_step, j.max_step, j.min_step, j.capture_margin,
                                    j.lambda_end))
    out = {}
    for mode in ("precise", "fast"):
        plan = batch.BatchPlan(tjobs, mode=mode)
        plan.launch()
        r = plan.results(as_numpy=True)
        out[mode] = (np.concatenate([x.states for x in r]), np.concatenate([x.status for x in r]))
        del plan
    np.testing.assert_array_equal(out["fast"][1], out["precise"][1])
    ok = out["precise"][1] == 6
    d = block_diff(out["fast"][0][ok], out["precise"][0][ok])
    print(f"fast vs precise over {int(ok.sum())} orbits: median {np.median(d):.2e}, "
          f"max {d.max():.2e}")
    assert float(d.max()) <= STATE_RTOL
    g = golden("c5_orbits.npz")
    k = 0
    while f"o{k}_index" in g.files:
        i = int(g[f"o{k}_index"])
        assert float(block_diff(out["fast"][0][i:i + 1], g[f"o{k}_final"])[0]) <= STATE_RTOL
        k += 1
