"""The quad tracer's right-hand side (csrc/bh_quad.inc) evaluates the three
gradient rows of the precise build's geodesic_rhs as one formula with
per-row operands, relying on exact IEEE identities (v + (-0.0) = v,
v * 1.0 = v, -0.0 - v = -v, a - b = a + (-b)) to keep each row's rounding.
Here those row formulas, operand for operand, are evaluated with NumPy and
compared bit for bit with the oracle's right-hand side (geodesics.py:42-74
in NumPy's order) on random states that include signed-zero coordinates
and momenta.  A dropped identity (e.g. +0.0 for -0.0) or a reordered
product shows up as a changed bit pattern.  The GPU tests compare the
kernel itself with the one-lane kernels."""

import numpy as np
import pytest

from oracle import bh_oracle as orc

TINY = 1e-30


def _quad_rows(st, q, x, y, z, pt, px, py, pz, zero=-0.0):
    """(k[1 + q], k[5 + q], k[0]) from the quad's uniform row formula;
    ``zero`` is the additive identity used for absent terms (-0.0 is exact
    for every operand, +0.0 is not)."""
    m, a = st.mass, st.spin
    a2 = a * a
    four_a2, three_a2, dhz_c = 4.0 * a2, 3.0 * a2, -2.0 * m * a2
    xq, pq = (x, y, z)[q], (px, py, pz)[q]
    row01, row0, row1 = q < 2, q == 0, q == 1
    gam = 1.0 if q == 2 else zero
    rho2 = x * x + y * y + z * z
    qq = rho2 - a2
    s = np.maximum(np.sqrt(qq * qq + four_a2 * z * z), TINY)
    r2 = 0.5 * (qq + s)
    r = np.sqrt(np.maximum(r2, TINY))
    r2a2 = r2 + a2
    qd = np.maximum(r2 * r2 + a2 * z * z, TINY)
    h = (m * r * r2) / qd
    inv_r2a2 = 1.0 / np.maximum(r2a2, TINY)
    inv_r = 1.0 / r
    inv_s = 1.0 / s
    lx = (r * x + a * y) * inv_r2a2
    ly = (r * y - a * x) * inv_r2a2
    lz = z * inv_r
    rq = xq * (r if row01 else r2a2) * (1.0 if row01 else inv_r) * inv_s
    inv_qd2 = 1.0 / (qd * qd)
    dh_dr = m * r2 * (three_a2 * z * z - r2 * r2) * inv_qd2
    dh_dz = dhz_c * z * r * r2 * inv_qd2
    ghq = dh_dr * rq + (zero if row01 else dh_dz)
    two_r = 2.0 * r
    al = r if row0 else (a if row1 else zero)
    be = r if row1 else (-a if row0 else zero)
    g0 = (rq * x + al) * inv_r2a2 - lx * two_r * rq * inv_r2a2
    g1 = (rq * y + be) * inv_r2a2 - ly * two_r * rq * inv_r2a2
    g2 = (gam - lz * rq) * inv_r
    # the kernel starts each three-term sum from +0.0, as NumPy's reduction
    lp = -pt + (((0.0 + lx * px) + ly * py) + lz * pz)
    two_h_lp = 2.0 * h * lp
    cq = ((0.0 + g0 * px) + g1 * py) + g2 * pz
    lq = (lx, ly, lz)[q]
    return pq - two_h_lp * lq, ghq * (lp * lp) + two_h_lp * cq, -pt + two_h_lp


def _states(rng, n):
    s = rng.normal(size=(n, 8)) * rng.choice([1e-3, 1.0, 10.0, 300.0], size=(n, 1))
    s[::7, 3] = 0.0
    s[::11, 1] = -0.0
    s[::13, 5] = 0.0
    s[::17, 2] = -0.0
    s[::19, 1:4] = 0.0
    s[::23, 6] = -0.0
    return s


@pytest.mark.parametrize("spin", [0.0, 0.5, 0.9, -0.998, 0.998])
def test_quad_row_formulas_equal_the_reference_rhs_bitwise(spin):
    st = orc.Kerr(spin=spin)
    s = _states(np.random.default_rng(int(1000 * abs(spin)) + (spin < 0)), 40000)
    with np.errstate(all="ignore"):
        ref = orc.geodesic_rhs(st, s)
        cols = tuple(s[:, i] for i in (1, 2, 3, 4, 5, 6, 7))
        for q in range(3):
            kx, kp, k0 = _quad_rows(st, q, *cols)
            for got, want in ((kx, ref[:, 1 + q]), (kp, ref[:, 5 + q]), (k0, ref[:, 0])):
                both_nan = np.isnan(got) & np.isnan(want)
                same = got.view(np.int64) == want.view(np.int64)
                assert np.all(same | both_nan), (spin, q)


def _mismatches(spin, zero):
    st = orc.Kerr(spin=spin)
    s = _states(np.random.default_rng(7), 20000)
    bad = 0
    with np.errstate(all="ignore"):
        ref = orc.geodesic_rhs(st, s)
        cols = tuple(s[:, i] for i in (1, 2, 3, 4, 5, 6, 7))
        for q in range(3):
            kx, kp, _ = _quad_rows(st, q, *cols, zero=zero)
            for got, want in ((kx, ref[:, 1 + q]), (kp, ref[:, 5 + q])):
                both_nan = np.isnan(got) & np.isnan(want)
                bad += int(np.sum((got.view(np.int64) != want.view(np.int64)) & ~both_nan))
    return bad


def test_a_dropped_identity_is_detected():
    """With +0.0 as the absent term (not an identity for -0.0 operands) the
    same comparison finds changed bits (signed zeros of the momentum
    derivative), so the bitwise check above has teeth."""
    assert _mismatches(0.0, -0.0) == 0
    assert _mismatches(0.0, 0.0) > 0
