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


def assert_trace_parity(got_states, got_status, got_steps, ref_states, ref_status, ref_steps,
                        allow_mismatch=0, label=""):
    """Status and accepted-step counts identical, final states within
    STATE_RTOL; up to ``allow_mismatch`` rays may differ (rays on the
    critical curve, where round-off decides the outcome)."""
    st_ok = np.asarray(got_status) == np.asarray(ref_status)
    steps_ok = np.asarray(got_steps) == np.asarray(ref_steps)
    s_ok = state_close(got_states, ref_states)
    bad = ~(st_ok & steps_ok & s_ok)
    nbad = int(bad.sum())
    if nbad > allow_mismatch:
        idx = np.nonzero(bad)[0][:10]
        raise AssertionError(
            f"{label}: {nbad} rays differ (allowed {allow_mismatch}); first {idx.tolist()}; "
            f"status {np.asarray(got_status)[idx].tolist()} vs {np.asarray(ref_status)[idx].tolist()}; "
            f"steps {np.asarray(got_steps)[idx].tolist()} vs {np.asarray(ref_steps)[idx].tolist()}; "
            f"max rel state diff over agreeing rays "
            f"{max_rel_diff(np.asarray(got_states)[~bad], np.asarray(ref_states)[~bad]):.3e}")
    return nbad
