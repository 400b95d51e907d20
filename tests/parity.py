"""Comparison helpers shared by the oracle and GPU parity tests."""

import numpy as np

STATE_RTOL = 1e-7   # final positions and momenta: |d| <= 1e-7 max(|ref|, 1)


def state_close(got, ref, rtol=STATE_RTOL):
    """Per-ray mask: every component within rtol * max(|ref|, 1)."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    both_nan = np.isnan(got) & np.isnan(ref)
    same_inf = np.isinf(got) & np.isinf(ref) & (np.sign(got) == np.sign(ref))
    ok = (np.abs(got - ref) <= rtol * np.maximum(np.abs(ref), 1.0)) | both_nan | same_inf
    return np.all(ok, axis=-1)


def max_rel_diff(got, ref):
    got = np.asarray(got)
    ref = np.asarray(ref)
    fin = np.isfinite(got) & np.isfinite(ref)
    d = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    return float(np.max(np.where(fin, d, 0.0))) if d.size else 0.0


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
