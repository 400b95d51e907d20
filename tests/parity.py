"""Comparison helpers shared by the oracle and GPU parity tests."""

import numpy as np

STATE_RTOL = 1e-7   # final positions and momenta: |d| <= 1e-7 max(|ref|, 1)


def _blocks(a):
    """Split (n, 8) states into the four physical blocks: t, position
    vector, p_t, momentum vector."""
    a = np.asarray(a)
    return [a[:, 0:1], a[:, 1:4], a[:, 4:5], a[:, 5:8]]


def state_close(got, ref, rtol=STATE_RTOL):
    """Per-ray mask: t, x, p_t and p each within rtol of the reference,
    relative to max(|block|, 1) with |.| the Euclidean norm of the block.

    Norm-relative (not per-component) because a captured ray ends just
    inside the horizon with covariant momenta of order 1e3 whose small
    components carry the absolute round-off of the large ones; a 1-ulp
    change of one step factor moves such a component by ~1e-7 of itself
    in the reference's own NumPy arithmetic.
    """
    ok = np.ones(np.asarray(ref).shape[0], dtype=bool)
    for g, r in zip(_blocks(got), _blocks(ref)):
        with np.errstate(invalid="ignore"):
            diff = np.linalg.norm(g - r, axis=-1)
        scale = np.maximum(np.linalg.norm(r, axis=-1), 1.0)
        same_nonfinite = np.all((np.isnan(g) & np.isnan(r)) | (g == r), axis=-1)
        ok &= (diff <= rtol * scale) | same_nonfinite
    return ok


def max_rel_diff(got, ref):
    """Largest block-norm relative difference (see state_close)."""
    worst = 0.0
    for g, r in zip(_blocks(got), _blocks(ref)):
        if g.shape[0] == 0:
            continue
        with np.errstate(invalid="ignore"):
            d = np.linalg.norm(g - r, axis=-1) / np.maximum(np.linalg.norm(r, axis=-1), 1.0)
        d = d[np.isfinite(d)]
        if d.size:
            worst = max(worst, float(d.max()))
    return worst


def block_diff(got, ref):
    """Per-ray largest block-norm relative difference (see state_close)."""
    out = np.zeros(np.asarray(ref).shape[0])
    for g, r in zip(_blocks(got), _blocks(ref)):
        with np.errstate(invalid="ignore"):
            d = np.linalg.norm(g - r, axis=-1) / np.maximum(np.linalg.norm(r, axis=-1), 1.0)
        same = np.all((np.isnan(g) & np.isnan(r)) | (g == r), axis=-1)
        out = np.maximum(out, np.where(same, 0.0, np.nan_to_num(d, nan=np.inf)))
    return out


def reference_spread(spin, state0, ref_states, **trace_kw):
    """How far the reference's own final states move under ulp-level changes
    of its step factors: the largest block difference over three re-traces
    by the oracle (which reproduces the reference bit for bit on this host),
    one with libm pow in place of NumPy's power ufunc (a different but
    equally valid host) and two with the pow result moved by a random ulp."""
    from oracle import bh_oracle as orc
    import warnings

    worst = np.zeros(np.asarray(state0).shape[0])
    for fn in (orc.step_factor_libm, orc.step_factor_ulp_noise(1),
               orc.step_factor_ulp_noise(2)):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            alt = orc.trace_rays(orc.Kerr(spin=spin), state0, factor_fn=fn, **trace_kw)
        worst = np.maximum(worst, block_diff(alt.states, ref_states))
    return worst


SPREAD_FACTOR = 10.0
# rays allowed to meet the state bar only through the spread allowance: at
# most this fraction of a fixture (at least one ray), all ending CAPTURED
SPREAD_ONLY_FRACTION = 0.01


def assert_trace_parity(got_states, got_status, got_steps, ref_states, ref_status, ref_steps,
                        allow_mismatch=0, label="", spread=None, max_spread_only=None,
                        spread_status=(1,)):
    """Status and accepted-step counts identical; final states within
    STATE_RTOL, or within SPREAD_FACTOR x the reference's own host-to-host
    spread on rays where that spread exceeds STATE_RTOL (captured rays that
    end inside the horizon with momenta of order 1e3 and shrinking steps are
    that ill-conditioned); up to ``allow_mismatch`` rays may differ.

    The rays that meet the state bar only through that allowance are
    counted and printed with the worst of their differences (in units of
    STATE_RTOL); at most ``max_spread_only`` of them (default 1% of the
    rays, at least one) are allowed, and each must end with a status in
    ``spread_status`` (CAPTURED).  Returns (mismatches, spread-only rays,
    worst spread-only difference / STATE_RTOL)."""
    ref_status = np.asarray(ref_status)
    n = ref_status.shape[0]
    st_ok = np.asarray(got_status) == ref_status
    steps_ok = np.asarray(got_steps) == np.asarray(ref_steps)
    tol = np.full(n, STATE_RTOL)
    if spread is not None:
        tol = np.maximum(tol, SPREAD_FACTOR * np.asarray(spread))
    d = block_diff(got_states, ref_states)
    s_ok = d <= tol
    bad = ~(st_ok & steps_ok & s_ok)
    nbad = int(bad.sum())
    via = ~bad & (d > STATE_RTOL)
    n_via = int(via.sum())
    worst = float(np.max(d[via], initial=0.0)) / STATE_RTOL
    print(f"[parity] {label}: {n} rays, {nbad} mismatched, {n_via} within the state bar only "
          f"through the reference's spread (worst {worst:.2f} x {STATE_RTOL:g}, statuses "
          f"{sorted(set(ref_status[via].tolist()))})")
    if nbad > allow_mismatch:
        idx = np.nonzero(bad)[0][:10]
        raise AssertionError(
            f"{label}: {nbad} rays differ (allowed {allow_mismatch}); first {idx.tolist()}; "
            f"status {np.asarray(got_status)[idx].tolist()} vs {ref_status[idx].tolist()}; "
            f"steps {np.asarray(got_steps)[idx].tolist()} vs {np.asarray(ref_steps)[idx].tolist()}; "
            f"state diff {d[idx].tolist()} tol {tol[idx].tolist()}; "
            f"max diff over agreeing rays {float(np.max(d[~bad], initial=0.0)):.3e}")
    cap = max_spread_only if max_spread_only is not None else max(1, int(SPREAD_ONLY_FRACTION * n))
    assert n_via <= cap, f"{label}: {n_via} rays pass only through the spread allowance (cap {cap})"
    assert np.all(np.isin(ref_status[via], spread_status)), \
        f"{label}: spread-only rays with statuses {ref_status[via].tolist()}"
    return nbad, n_via, worst
