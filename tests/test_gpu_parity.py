"""GPU parity: the HIP kernels (through the C ABI) against the oracle and the
reference's golden vectors.

Bars: precise build -- geometry, RHS, DP5 stages and initial conditions equal
the oracle bit for bit (every operation correctly rounded, NumPy order);
traces -- identical status, accepted steps and iterations, final states
within 1e-7 relative (t, position, p_t and momentum blocks, each relative to
max(norm, 1)); on rays where the reference itself moves by more than that
between hosts (NumPy's SIMD pow vs libm pow: captured rays ending inside the
horizon at tight rtol), within 10x that host-to-host spread.  The precise
build meets this at every tolerance.  The fast build rounds the RHS
differently in the last ulp; the embedded error estimate amplifies that by
~1/rtol and the step sizes follow it, so it is held to the bar at
rtol >= 1e-8 and bounded separately below.  mode="auto" (the default) picks
fast at rtol >= 1e-8 and precise below.
"""

import math
import warnings

import numpy as np
import pytest

from conftest import golden
from oracle import bh_oracle as orc
from parity import assert_trace_parity, max_rel_diff, reference_spread
from blackhorizon_amd import workloads

pytestmark = pytest.mark.gpu

SPINS = [0.0, 0.5, 0.9, -0.998]
MODES = ["precise", "fast", "auto"]
FAST_MIN_RTOL = 1e-8


def mode_applies(mode, rtol):
    return not (mode == "fast" and rtol < FAST_MIN_RTOL)


def points(n, seed, r_min=1.2, r_max=80.0):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    return rng.uniform(r_min, r_max, size=(n, 1)) * u


def states(spin, n, seed):
    st = orc.Kerr(spin=spin)
    pos = points(n, seed, 2.5, 40.0)
    d = np.random.default_rng(seed + 1).normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, d, "past"))


@pytest.mark.parametrize("spin", SPINS)
def test_geometry_precise_bit_exact(hip, spin):
    p = points(4096, 1)
    got = hip.KerrSpacetime(spin=spin).geometry(p[:, 0], p[:, 1], p[:, 2], gradients=True,
                                                mode="precise")
    ref = orc.Kerr(spin=spin).geometry(p[:, 0], p[:, 1], p[:, 2], gradients=True)
    for f in ("radius", "h", "l", "grad_h", "grad_l"):
        np.testing.assert_array_equal(getattr(got, f), getattr(ref, f), err_msg=f)


@pytest.mark.parametrize("spin", SPINS)
def test_geometry_fast_close(hip, spin):
    p = points(4096, 2)
    got = hip.KerrSpacetime(spin=spin).geometry(p[:, 0], p[:, 1], p[:, 2], gradients=True,
                                                mode="fast")
    ref = orc.Kerr(spin=spin).geometry(p[:, 0], p[:, 1], p[:, 2], gradients=True)
    for f in ("radius", "h", "l", "grad_h", "grad_l"):
        a, b = getattr(got, f), getattr(ref, f)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14 * np.max(np.abs(b)), err_msg=f)


@pytest.mark.parametrize("spin", SPINS)
def test_rhs(hip, spin):
    from blackhorizon_amd import geodesics

    s = states(spin, 2048, 3)
    ref = orc.geodesic_rhs(orc.Kerr(spin=spin), s)
    st = hip.KerrSpacetime(spin=spin)
    np.testing.assert_array_equal(geodesics.geodesic_rhs(st, s, mode="precise"), ref)
    fast = geodesics.geodesic_rhs(st, s, mode="fast")
    np.testing.assert_allclose(fast, ref, rtol=1e-11, atol=1e-13)


def _bits_equal(a, b, label):
    a, b = np.asarray(a), np.asarray(b)
    same = a.view(np.int64) == b.view(np.int64)
    assert np.all(same | (np.isnan(a) & np.isnan(b))), label


@pytest.mark.parametrize("spin", [0.0, 0.9, -0.998])
def test_precise_signed_zeros_bitwise(hip, spin):
    """States with exact +-0.0 coordinates and momenta (the origin, the
    axis, the equatorial plane, zero momentum components): the precise
    build's right-hand side, geometry and diagnostics equal the oracle's
    raw bit patterns, signed zeros included -- NumPy's three-term sums start
    from +0.0, so an all -0.0 sum is +0.0 there and on the device."""
    from blackhorizon_amd import geodesics

    rng = np.random.default_rng(17)
    s = rng.normal(size=(4096, 8)) * rng.choice([1e-3, 1.0, 10.0, 300.0], size=(4096, 1))
    s[::3, 3] = 0.0
    s[::5, 1] = -0.0
    s[::7, 2] = 0.0
    s[::11, 1:4] = 0.0
    s[::13, 5] = -0.0
    s[::17, 6] = -0.0
    s[::19, 7] = 0.0
    s[::23, 5:8] = -0.0
    ost, st = orc.Kerr(spin=spin), hip.KerrSpacetime(spin=spin)
    with np.errstate(all="ignore"):
        ref = orc.geodesic_rhs(ost, s)
        _bits_equal(geodesics.geodesic_rhs(st, s, mode="precise"), ref, "rhs")
        got = st.geometry(s[:, 1], s[:, 2], s[:, 3], gradients=True, mode="precise")
        want = ost.geometry(s[:, 1], s[:, 2], s[:, 3], gradients=True)
        for f in ("radius", "h", "l", "grad_h", "grad_l"):
            _bits_equal(getattr(got, f), getattr(want, f), f)
        _bits_equal(geodesics.hamiltonian(st, s), orc.hamiltonian(ost, s), "hamiltonian")
        _bits_equal(geodesics.coordinate_velocity(st, s), orc.coordinate_velocity(ost, s),
                    "coordinate_velocity")


@pytest.mark.parametrize("spin", [0.9, -0.5])
def test_dp5_step(hip, spin):
    from blackhorizon_amd import integrators

    s = states(spin, 2048, 4)
    h = np.random.default_rng(5).uniform(0.01, 2.0, 2048)
    ost = orc.Kerr(spin=spin)
    new_r, err_r = orc.dormand_prince_step(lambda y: orc.geodesic_rhs(ost, y), s, h)
    ratio_r = orc.error_ratio(s, new_r, err_r, 1e-8, 1e-10)
    fac_r = orc.step_factor(ratio_r)
    g = integrators.GeodesicRHS(hip.KerrSpacetime(spin=spin), mode="precise")
    new, err, ratio, fac = integrators.dormand_prince_attempt(g, s, h, 1e-8, 1e-10)
    np.testing.assert_array_equal(new, new_r)
    np.testing.assert_array_equal(err, err_r)
    np.testing.assert_array_equal(ratio, ratio_r)
    # x ** -0.2 is correctly rounded on the device; NumPy's power is off by an
    # ulp in ~5% of calls, so the factors agree to 1-2 ulp
    np.testing.assert_allclose(fac, fac_r, rtol=4e-16)
    # against the correctly rounded factor: x ** -0.2 in extended precision
    # (the double 0.2 is 3602879701896397 / 2^54) rounded once to double,
    # then 0.9 * that and the clip, as the reference computes them
    safe = np.maximum(ratio_r, 1e-16)
    e = np.longdouble(-3602879701896397) / np.longdouble(2.0 ** 54)
    pw = np.power(safe.astype(np.longdouble), e).astype(np.float64)
    fac_cr = np.clip(0.9 * pw, 0.2, 5.0)
    assert np.mean(fac == fac_cr) > 0.99
    np.testing.assert_allclose(fac, fac_cr, rtol=2.3e-16)
    g = integrators.GeodesicRHS(hip.KerrSpacetime(spin=spin), mode="fast")
    new, err, ratio, fac = integrators.dormand_prince_attempt(g, s, h, 1e-8, 1e-10)
    # FMA-contracted stages round differently; the embedded error is a sum
    # that cancels to O(h^5), so its relative rounding error is larger than
    # that of the state, and the ratio is compared on an absolute scale.
    np.testing.assert_allclose(new, new_r, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(ratio, ratio_r, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(fac, orc.step_factor(ratio), rtol=3e-15)


def _controller_rows(n, dim, seed):
    """Rows of dim values over ~17 decades of |y| and ~17 of |err|, with
    rows that hold a NaN, an infinity, a zero scale and an all-zero error."""
    rng = np.random.default_rng(seed)
    y = rng.normal(size=(n, dim)) * np.exp(rng.uniform(-20.0, 20.0, (n, dim)))
    yn = y + 1e-3 * np.abs(y) * rng.normal(size=(n, dim))
    err = rng.normal(size=(n, dim)) * np.exp(rng.uniform(-40.0, 0.0, (n, dim)))
    err[0, 0] = np.nan
    err[1, dim - 1] = np.inf
    y[2, :] = yn[2, :] = 0.0  # scale = atol = 0 below: 0/0 and x/0
    err[3, :] = 0.0
    return y, yn, err


@pytest.mark.parametrize("dim", [1, 2, 7, 8, 9, 16, 23, 128])
def test_error_ratio_any_row_width(hip, dim):
    """error_ratio (integrators.py:110-117) over rows of 1..128 values: the
    precise build equals NumPy bit for bit (division, squares, NumPy's
    summation order for that width, mean); non-finite rows give +inf; the
    fast build agrees to 1e-14 relative."""
    import torch

    from blackhorizon_amd import integrators

    y, yn, err = _controller_rows(1000, dim, 100 + dim)
    for atol in (1e-12, 0.0):
        with np.errstate(divide="ignore"):  # atol 0 and a zero state: scale 0, ratio inf
            want = orc.error_ratio(y, yn, err, 1e-9, atol)
        got = integrators.error_ratio(y, yn, err, 1e-9, atol)
        np.testing.assert_array_equal(got, want)
        fast = integrators.error_ratio(y, yn, err, 1e-9, atol, mode="fast")
        np.testing.assert_array_equal(np.isinf(fast), np.isinf(want))
        np.testing.assert_allclose(fast, want, rtol=1e-14)
    assert np.isinf(want[:3]).all() and want[3] == 0.0
    # leading axes and tensors: (4, 250, dim) -> (4, 250); one row -> 0-d
    t = [torch.from_numpy(a.reshape(4, 250, dim)).cuda() for a in (y, yn, err)]
    got_t = integrators.error_ratio(*t, 1e-9, 1e-12)
    assert isinstance(got_t, torch.Tensor) and got_t.shape == (4, 250)
    np.testing.assert_array_equal(got_t.cpu().numpy().reshape(-1),
                                  orc.error_ratio(y, yn, err, 1e-9, 1e-12))
    one = integrators.error_ratio(y[5], yn[5], err[5], 1e-9, 1e-12)
    assert np.shape(one) == ()
    assert one == orc.error_ratio(y[5:6], yn[5:6], err[5:6], 1e-9, 1e-12)[0]


def test_error_ratio_reference_nan_case(hip):
    """The reference's own case (tests/test_integrators.py:71-75): a NaN in
    a two-value row gives an infinite ratio."""
    from blackhorizon_amd import integrators

    state = np.array([[1.0, 0.0]])
    ratio = integrators.error_ratio(state, state, np.array([[np.nan, 0.0]]), rtol=1e-9,
                                    atol=1e-12)
    assert np.isinf(ratio[0])


def test_step_factor_api(hip):
    """step_factor (integrators.py:120-124) over 100 decades of ratios and the
    special values: the precise build's power is correctly rounded (against
    x ** -0.2 in extended precision, rounded once), so it equals that
    reference in > 99% of values and is within an ulp everywhere; NaN stays
    NaN, 0, negative and tiny ratios give 5, +inf gives 0.2.  The fast build
    is within 3e-15 relative."""
    import torch

    from blackhorizon_amd import integrators

    rng = np.random.default_rng(21)
    special = np.array([0.0, -1.0, -np.inf, 1e-16, 1e-17, 1e-300, np.inf, np.nan, 1.0,
                        32.0 ** 5 / 1e5, 1e6, 1e-6, 0.9 ** 5 * 3125.0, 5e-324])
    ratio = np.concatenate([np.exp(rng.uniform(-115.0, 115.0, 20000)), special])
    want = orc.step_factor(ratio)
    safe = np.maximum(ratio, 1e-16)
    e = np.longdouble(-3602879701896397) / np.longdouble(2.0 ** 54)
    with np.errstate(invalid="ignore"):
        pw = np.power(safe.astype(np.longdouble), e).astype(np.float64)
    want_cr = np.clip(0.9 * pw, 0.2, 5.0)
    got = integrators.step_factor(ratio)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ratio))
    ok = ~np.isnan(ratio)
    assert np.mean(got[ok] == want_cr[ok]) > 0.99
    np.testing.assert_allclose(got[ok], want_cr[ok], rtol=2.3e-16)
    np.testing.assert_allclose(got, want, rtol=4e-16)
    np.testing.assert_array_equal(got[-14:-7], [5.0, 5.0, 5.0, 5.0, 5.0, 5.0, 0.2])
    fast = integrators.step_factor(ratio, mode="fast")
    np.testing.assert_allclose(fast, want, rtol=3e-15)
    t = torch.from_numpy(ratio[:20000].reshape(100, 200)).cuda()
    got_t = integrators.step_factor(t)
    assert isinstance(got_t, torch.Tensor) and got_t.shape == (100, 200)
    np.testing.assert_array_equal(got_t.cpu().numpy().reshape(-1), got[:20000])


@pytest.mark.parametrize("spin", SPINS)
def test_initial_conditions_and_diagnostics(hip, spin):
    from blackhorizon_amd import geodesics

    st = hip.KerrSpacetime(spin=spin)
    ost = orc.Kerr(spin=spin)
    pos = points(1024, 6, 3.0, 60.0)
    d = np.random.default_rng(7).normal(size=(1024, 3))
    for orient in ("past", "future"):
        for norm in (True, False):
            got = geodesics.null_momentum_from_velocity(st, pos, d, orient, norm)
            ref = orc.null_momentum_from_velocity(ost, pos, d, orient, norm)
            np.testing.assert_array_equal(got, ref)
    s = orc.build_state(pos, orc.null_momentum_from_velocity(ost, pos, d, "past"))
    np.testing.assert_array_equal(geodesics.hamiltonian(st, s), orc.hamiltonian(ost, s))
    np.testing.assert_array_equal(geodesics.coordinate_velocity(st, s),
                                  orc.coordinate_velocity(ost, s))
    e, lz = geodesics.conserved_quantities(s, st)
    er, lzr = orc.conserved_quantities(s)
    np.testing.assert_array_equal(e, er)
    np.testing.assert_array_equal(lz, lzr)
    r = np.linalg.norm(pos, axis=-1)
    om = 1.0 / (r ** 1.5 + abs(spin) + 1.0)
    v = np.stack([-om * pos[:, 1], om * pos[:, 0], np.zeros(1024)], -1) * 0.5
    np.testing.assert_array_equal(geodesics.timelike_momentum(st, pos, v),
                                  orc.timelike_momentum(ost, pos, v))
    with pytest.raises(ValueError):
        geodesics.timelike_momentum(st, pos[:2], np.array([[0.0, 2.0, 0.0], [0.0, 0.1, 0.0]]))


EDGE_CASES = ["default", "nonfinite", "max_steps_1", "max_steps_7", "min_step_high", "loose",
              "tight_small_h", "max_steps_0"]
# the same outcomes at rtol 1e-8 (edge_cases_r8.npz), where the fast build
# runs: every mode is held to the bar
EDGE_CASES_R8 = ["default", "nonfinite", "nonfinite_pos", "max_steps_0", "max_steps_neg",
                 "max_steps_1", "max_steps_7", "attempt_cap", "min_step_high", "small_h", "huge_h"]


def _edge_params():
    """(fixture, case, mode) triples: every mode where it applies, so that no
    case is skipped; the fast build on the tight cases is bounded by
    test_fast_build_tight_tolerance_bounded."""
    g = golden("edge_cases.npz")
    out = []
    for case in EDGE_CASES:
        rtol = float(g[case + "_kw_rtol"]) if case + "_kw_rtol" in g.files else 1e-9
        out += [("edge_cases.npz", case, m) for m in MODES if mode_applies(m, rtol)]
    out += [("edge_cases_r8.npz", case, m) for case in EDGE_CASES_R8 for m in MODES]
    return out


def _kw(g, name):
    pre = name + "_kw_"
    kw = {k[len(pre):]: float(g[k]) for k in g.files if k.startswith(pre)}
    if "max_steps" in kw:
        kw["max_steps"] = int(kw["max_steps"])
    return kw


_SPREADS = {}


def _spread(key, spin, state0, ref_states, kw):
    """Cached host-to-host spread of the reference's final states."""
    if key not in _SPREADS:
        _SPREADS[key] = reference_spread(spin, state0, ref_states, **kw)
    return _SPREADS[key]


@pytest.mark.parametrize("fixture,case,mode", _edge_params())
def test_trace_edge_cases(hip, fixture, case, mode):
    """Failure and cap outcomes against the reference: NaN and inf rows
    (FAILED; the fast build's NaN-dropping max in the error norm must still
    end them FAILED), max_steps 0 / negative / 1 / 7 and the 4 max_steps
    attempt cap (MAX_STEPS), h below min_step (FAILED), tiny and huge steps,
    a ray starting inside the horizon.  Precise-build traces this small run
    on the quad tracer by default; with an explicit launch shape they run
    one lane per ray: both are checked."""
    from blackhorizon_amd.tracer import resolve_mode

    g = golden(fixture)
    kw = _kw(g, case)
    shapes = [None]
    if resolve_mode(mode, kw.get("rtol", 1e-9)) == "precise":
        shapes.append(hip.Tune())
    spread = _spread((fixture, case), 0.9, g[case + "_state0"], g[case + "_states"], kw)
    for tune in shapes:
        res = hip.trace_rays(hip.KerrSpacetime(spin=0.9), g[case + "_state0"], mode=mode,
                             tune=tune, **kw)
        label = f"{fixture}:{case}:{mode}:{'one-lane' if tune else 'default'}"
        assert res.iterations == int(g[case + "_iterations"]), label
        assert_trace_parity(res.states, res.status, res.steps, g[case + "_states"],
                            g[case + "_status"], g[case + "_steps"], label=label, spread=spread)
        assert np.all(res.attempts >= res.steps)
        assert int(np.max(res.attempts, initial=0)) == res.iterations


@pytest.mark.parametrize("case", [c for c in EDGE_CASES if c != "loose"])
def test_fast_build_tight_tolerance_bounded(hip, case):
    """At rtol 1e-9 to 1e-12 the fast build keeps every outcome and stays
    within the drift its last-ulp RHS rounding allows (1e-16 / rtol scale),
    while the precise and auto builds meet the 1e-7 bar
    (test_trace_edge_cases)."""
    g = golden("edge_cases.npz")
    res = hip.trace_rays(hip.KerrSpacetime(spin=0.9), g[case + "_state0"], mode="fast",
                         **_kw(g, case))
    np.testing.assert_array_equal(res.status, g[case + "_status"])
    assert max_rel_diff(res.states, g[case + "_states"]) < 1e-3


@pytest.mark.parametrize("mode", MODES)
def test_trace_golden_small_frame(hip, mode):
    g = golden("small_frame_stride8.npz")
    job = workloads.small_frame()
    res = hip.trace_rays(hip.KerrSpacetime(spin=job.spin), g["state0"], job.escape_radius,
                         max_steps=job.max_steps, rtol=job.rtol, atol=job.atol,
                         capture_margin=job.capture_margin, mode=mode)
    assert res.iterations == int(g["iterations"])
    kw = dict(escape_radius=job.escape_radius, max_steps=job.max_steps, rtol=job.rtol,
              atol=job.atol, capture_margin=job.capture_margin)
    spread = _spread("small8", job.spin, g["state0"], g["states"], kw)
    assert_trace_parity(res.states, res.status, res.steps, g["states"], g["status"], g["steps"],
                        label="small frame", spread=spread)


@pytest.mark.parametrize("mode", MODES)
def test_trace_golden_fans_uneven_hd(hip, mode):
    g = golden("fans_d30.npz")
    for spin in (0.9, 0.998):
        k = f"a{spin}"
        if not mode_applies(mode, 1e-10):
            break
        res = hip.trace_rays(hip.KerrSpacetime(spin=spin), g[k + "_state0"], 60.0, rtol=1e-10,
                             mode=mode)
        assert res.iterations == int(g[k + "_iterations"])
        spread = _spread(("fan", k), spin, g[k + "_state0"], g[k + "_states"],
                         dict(escape_radius=60.0, rtol=1e-10))
        assert_trace_parity(res.states, res.status, res.steps, g[k + "_states"],
                            g[k + "_status"], g[k + "_steps"], label=k, spread=spread)
    g = golden("uneven_subsample.npz")
    for j, job in enumerate(workloads.uneven_mix()[:8]):
        if not mode_applies(mode, job.rtol):
            continue
        res = hip.trace_rays(hip.KerrSpacetime(spin=job.spin), g[f"j{j}_state0"],
                             job.escape_radius, max_steps=job.max_steps, rtol=job.rtol,
                             atol=job.atol, capture_margin=job.capture_margin, mode=mode)
        assert res.iterations == int(g[f"j{j}_iterations"])
        kw = dict(escape_radius=job.escape_radius, max_steps=job.max_steps, rtol=job.rtol,
                  atol=job.atol, capture_margin=job.capture_margin)
        spread = _spread(("job", j), job.spin, g[f"j{j}_state0"], g[f"j{j}_states"], kw)
        assert_trace_parity(res.states, res.status, res.steps, g[f"j{j}_states"],
                            g[f"j{j}_status"], g[f"j{j}_steps"], label=f"job {j}", spread=spread)
    g = golden("hd_subsample.npz")
    for f, job in enumerate(workloads.hd_frames()):
        res = hip.trace_rays(hip.KerrSpacetime(spin=job.spin), g[f"f{f}_state0"],
                             job.escape_radius, max_steps=job.max_steps, rtol=job.rtol,
                             atol=job.atol, capture_margin=job.capture_margin, mode=mode)
        kw = dict(escape_radius=job.escape_radius, max_steps=job.max_steps, rtol=job.rtol,
                  atol=job.atol, capture_margin=job.capture_margin)
        spread = _spread(("hd", f), job.spin, g[f"f{f}_state0"], g[f"f{f}_states"], kw)
        assert_trace_parity(res.states, res.status, res.steps, g[f"f{f}_states"],
                            g[f"f{f}_status"], g[f"f{f}_steps"], label=f"hd frame {f}",
                            spread=spread)


@pytest.fixture(scope="module")
def small_frame_oracle():
    job = workloads.small_frame()
    st = orc.Kerr(spin=job.spin)
    cam = orc.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees, width=job.width,
                                       height=job.height)
    s0 = orc.camera_state(st, cam)
    kw = dict(escape_radius=job.escape_radius, max_steps=job.max_steps, rtol=job.rtol,
              atol=job.atol, capture_margin=job.capture_margin)
    res, att = orc.trace_rays(st, s0, return_attempts=True, **kw)
    spread = reference_spread(job.spin, s0, res.states, **kw)
    return job, s0, res, att, spread


@pytest.mark.parametrize("mode", MODES)
def test_trace_full_small_frame_vs_oracle(hip, small_frame_oracle, mode):
    """All 19,200 rays of the 160x120 a = 0.9 frame against the oracle,
    including the E, L_z and Hamiltonian drift bound."""
    job, s0, ref, ref_att, spread = small_frame_oracle
    st = hip.KerrSpacetime(spin=job.spin)
    res = hip.trace_rays(st, s0, job.escape_radius, max_steps=job.max_steps, rtol=job.rtol,
                         atol=job.atol, capture_margin=job.capture_margin, mode=mode)
    assert res.iterations == ref.iterations
    assert_trace_parity(res.states, res.status, res.steps, ref.states, ref.status, ref.steps,
                        label=f"small frame {mode}", spread=spread)
    np.testing.assert_array_equal(res.attempts, ref_att)
    # drift no worse than the reference's: E exact, L_z and H_ham within the
    # oracle's drift plus round-off
    ost = orc.Kerr(spin=job.spin)
    e0, lz0 = orc.conserved_quantities(s0)
    e1, lz1 = orc.conserved_quantities(res.states)
    er, lzr = orc.conserved_quantities(ref.states)
    np.testing.assert_array_equal(e1, e0)
    fin = np.isfinite(lzr)
    assert np.max(np.abs(lz1 - lz0)[fin]) <= np.max(np.abs(lzr - lz0)[fin]) * 1.01 + 1e-12
    h1 = np.abs(orc.hamiltonian(ost, res.states))
    hr = np.abs(orc.hamiltonian(ost, ref.states))
    assert np.max(h1[fin]) <= np.max(hr[fin]) * 1.01 + 1e-12


def test_trace_tensor_io_and_no_mutation(hip):
    import torch

    job = workloads.small_frame()
    g = golden("small_frame_stride8.npz")
    s0 = torch.from_numpy(g["state0"][:256].copy()).cuda()
    keep = s0.clone()
    res = hip.trace_rays(hip.KerrSpacetime(spin=job.spin), s0, job.escape_radius, rtol=job.rtol,
                         atol=job.atol, capture_margin=job.capture_margin)
    assert isinstance(res.states, torch.Tensor) and res.states.is_cuda
    assert res.status.dtype == torch.int32 and res.steps.dtype == torch.int64
    assert torch.equal(s0, keep)
    np.testing.assert_array_equal(res.status.cpu().numpy(), g["status"][:256])


def test_numpy_results_are_independent_copies(hip):
    """NumPy callers get their results through pinned host memory reused
    from torch's cache: they equal the tensor path's bit for bit, are
    ordinary writable NumPy arrays, and stay unchanged when later calls
    trace other rays (a reused block would overwrite them)."""
    import torch

    job = workloads.small_frame()
    g = golden("small_frame_stride8.npz")
    st = hip.KerrSpacetime(spin=job.spin)
    kw = dict(rtol=job.rtol, atol=job.atol, capture_margin=job.capture_margin)
    first = hip.trace_rays(st, g["state0"][:512], job.escape_radius, **kw)
    keep = [first.states.copy(), first.status.copy(), first.steps.copy()]
    dev = hip.trace_rays(st, torch.from_numpy(g["state0"][:512].copy()).cuda(),
                         job.escape_radius, **kw)
    np.testing.assert_array_equal(first.states, dev.states.cpu().numpy())
    np.testing.assert_array_equal(first.status, dev.status.cpu().numpy())
    np.testing.assert_array_equal(first.steps, dev.steps.cpu().numpy())
    for _ in range(3):  # same sizes: the likeliest to be handed the same blocks
        later = hip.trace_rays(st, g["state0"][512:1024], job.escape_radius, **kw)
        del later
    other = hip.trace_rays(st, g["state0"][512:1024], job.escape_radius, **kw)
    np.testing.assert_array_equal(first.states, keep[0])
    np.testing.assert_array_equal(first.status, keep[1])
    np.testing.assert_array_equal(first.steps, keep[2])
    assert not np.array_equal(other.states, first.states)
    for a in (first.states, first.status, first.steps):
        assert isinstance(a, np.ndarray) and a.flags.writeable
    first.states[0, 0] = -1.0  # writable without touching the other results
    assert other.states[0, 0] != -1.0


def test_trace_value_errors(hip):
    st = hip.KerrSpacetime(spin=0.0)
    pos = np.array([[50.0, 0.0, 0.0]])
    from blackhorizon_amd import geodesics

    mom = geodesics.null_momentum_from_velocity(st, pos, np.array([[-1.0, 0.0, 0.0]]))
    s0 = geodesics.build_state(pos, mom)
    with pytest.raises(ValueError):
        hip.trace_rays(st, s0, escape_radius=40.0)
    res = hip.trace_rays(st, np.zeros((0, 8)), escape_radius=40.0)
    assert res.states.shape == (0, 8) and res.iterations == 0


@pytest.mark.parametrize("tune", [(256, 1), (64, 1), (128, 3), (256, 8)])
def test_launch_shapes_bitwise_identical(hip, tune):
    g = golden("small_frame_stride8.npz")
    job = workloads.small_frame()
    kw = dict(max_steps=job.max_steps, rtol=job.rtol, atol=job.atol,
              capture_margin=job.capture_margin)
    st = hip.KerrSpacetime(spin=job.spin)
    base = hip.trace_rays(st, g["state0"], job.escape_radius, **kw)
    other = hip.trace_rays(st, g["state0"], job.escape_radius,
                           tune=hip.Tune(block_size=tune[0], rays_per_lane=tune[1]), **kw)
    np.testing.assert_array_equal(other.states, base.states)
    np.testing.assert_array_equal(other.status, base.status)
    np.testing.assert_array_equal(other.steps, base.steps)
    assert other.iterations == base.iterations


# --- shadow-boundary closed forms (reference test_shadow.py) ----------------

def _fan_states(spin, offsets, distance=1000.0):
    st = orc.Kerr(spin=spin)
    n = offsets.shape[0]
    pos = np.tile([[distance, 0.0, 0.0]], (n, 1))
    d = np.stack([-np.ones(n), offsets, np.zeros(n)], axis=-1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return orc.build_state(pos, orc.null_momentum_from_velocity(st, pos, d, "past"))


def _boundary(hip, spin, lo, hi, rounds=3, n=48, mode="fast"):
    st = hip.KerrSpacetime(spin=spin)
    b = None
    for _ in range(rounds):
        off = np.linspace(lo, hi, n)
        s0 = _fan_states(spin, off)
        res = hip.trace_rays(st, s0, 1200.0, max_steps=60000, rtol=1e-10, atol=1e-13,
                             max_step=5.0, mode=mode)
        e, lz = orc.conserved_quantities(s0)
        imp = lz / e
        cap = res.status == 1
        esc = res.status == 2
        assert cap.any() and esc.any() and (cap | esc).all()
        mag = np.abs(off)
        ic = int(np.argmax(np.where(cap, mag, -1.0)))
        lo = float(off[ic])
        ie = int(np.argmin(np.where(esc & (mag > mag[ic]), mag, np.inf)))
        hi = float(off[ie])
        b = 0.5 * (abs(imp[ic]) + abs(imp[ie]))
    return b


def _bardeen(spin, prograde):
    r = orc.Kerr(spin=spin).photon_orbit_radius(prograde)
    return abs((r * r * (r - 3.0) + spin * spin * (r + 1.0)) / (spin * (1.0 - r)))


def test_shadow_schwarzschild_critical_b(hip):
    b = _boundary(hip, 0.0, 4.5e-3, 6.5e-3)
    assert abs(b - 3.0 * math.sqrt(3.0)) / (3.0 * math.sqrt(3.0)) < 5e-3


def test_shadow_kerr_prograde_retrograde(hip):
    bp = _boundary(hip, 0.9, -4.0e-3, -2.0e-3)
    br = _boundary(hip, 0.9, 5.5e-3, 8.0e-3)
    assert abs(bp - _bardeen(0.9, True)) / _bardeen(0.9, True) < 5e-3
    assert abs(br - _bardeen(0.9, False)) / _bardeen(0.9, False) < 5e-3
    assert bp < br


def test_shadow_zero_spin_symmetric(hip):
    bp = _boundary(hip, 0.0, 4.5e-3, 6.5e-3, rounds=2)
    bm = _boundary(hip, 0.0, -6.5e-3, -4.5e-3, rounds=2)
    assert abs(bp - bm) / bp < 1e-3


def test_radial_infall_and_escape(hip):
    from blackhorizon_amd import geodesics

    st = hip.KerrSpacetime(spin=0.9)
    pos = np.array([[20.0, 0.0, 0.0], [20.0, 0.0, 0.0]])
    mom = geodesics.null_momentum_from_velocity(st, pos, np.array([[-1.0, 0, 0], [1.0, 0, 0]]),
                                                "future")
    res = hip.trace_rays(st, geodesics.build_state(pos, mom), escape_radius=30.0)
    assert res.status.tolist() == [1, 2]


def test_shadow_gpu_png_full_frame(hip):
    """Every pixel of the reference-held shadow_gpu.png (320x240, a = 0.9,
    inclination 85 deg, D = 60, FOV 20 deg): capture mask identical, escaped
    checker cell identical for >= 99.9% of escaped pixels."""
    from blackhorizon_amd import geodesics

    cls = golden("shadow_gpu_classes.npz")["classes"].ravel()
    job = workloads.shadow_cli_frame()
    st = hip.KerrSpacetime(spin=job.spin)
    cam = hip.PinholeCamera.from_orbit(job.distance, job.inclination_degrees,
                                       fov_degrees=job.fov_degrees, width=job.width,
                                       height=job.height)
    pos = cam.ray_origins()
    mom = geodesics.null_momentum_from_velocity(st, pos, cam.ray_directions(), "past")
    res = hip.trace_rays(st, geodesics.build_state(pos, mom), job.escape_radius,
                         max_steps=job.max_steps, rtol=job.rtol)
    np.testing.assert_array_equal(res.status == 1, cls == 0)
    esc = res.status == 2
    np.testing.assert_array_equal(esc, cls != 0)
    light = orc.checker_parity(orc.Kerr(spin=job.spin), res.states[esc])
    assert np.mean(light == (cls[esc] == 2)) > 0.999


def test_geometry_and_metric_keep_the_callers_shapes(hip):
    """kerr.py:118-135 and 201-235 take broadcastable (x, y, z) and (..., 4)
    positions: a (4, 5) grid of x with a row of y and a scalar z, and a
    (2, 3, 4) block of positions, give results on those shapes, bit for bit
    the oracle's (precise build), as NumPy or as device tensors."""
    import torch

    rng = np.random.default_rng(12)
    x = rng.uniform(-20.0, 20.0, (4, 5))
    y = rng.uniform(-20.0, 20.0, 5)
    z = 3.5
    st = hip.KerrSpacetime(spin=0.8)
    ost = orc.Kerr(spin=0.8)
    got = st.geometry(x, y, z, gradients=True)
    xb, yb, zb = np.broadcast_arrays(x, y, np.float64(z))
    ref = ost.geometry(xb.ravel(), yb.ravel(), zb.ravel(), gradients=True)
    for f, tail in (("radius", ()), ("h", ()), ("l", (3,)), ("grad_h", (3,)),
                    ("grad_l", (3, 3))):
        a = getattr(got, f)
        assert a.shape == (4, 5) + tail, (f, a.shape)
        np.testing.assert_array_equal(a.reshape((-1,) + tail), getattr(ref, f), err_msg=f)
    r = st.kerr_schild_radius(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(),
                              torch.tensor(z, dtype=torch.float64).cuda())
    assert isinstance(r, torch.Tensor) and tuple(r.shape) == (4, 5)
    np.testing.assert_array_equal(r.cpu().numpy(), got.radius)
    pos = np.concatenate([np.zeros((2, 3, 1)), rng.uniform(-15.0, 15.0, (2, 3, 3))], -1)
    g = st.metric(pos)
    gi = st.inverse_metric(pos)
    assert g.shape == (2, 3, 4, 4) and gi.shape == (2, 3, 4, 4)
    flat = pos.reshape(-1, 4)
    np.testing.assert_array_equal(g.reshape(-1, 4, 4), ost.metric(flat))
    np.testing.assert_array_equal(gi.reshape(-1, 4, 4), ost.inverse_metric(flat))
    np.testing.assert_allclose(np.einsum("...ij,...jk->...ik", g, gi),
                               np.broadcast_to(np.eye(4), (2, 3, 4, 4)), atol=1e-12)
    one = st.metric(pos[0, 0])
    assert one.shape == (4, 4)
    np.testing.assert_array_equal(one, g[0, 0])
