"""Pins the oracle's DP5(4) tableau, step, error ratio and step factor
(re-derived from the reference's test_integrators.py plus the tableau's
order conditions)."""

from fractions import Fraction as F

import numpy as np

from oracle import bh_oracle as orc


def harmonic(state):
    d = np.empty_like(state)
    d[:, 0] = state[:, 1]
    d[:, 1] = -state[:, 0]
    return d


def _run(step, n_steps, t=2 * np.pi):
    s = np.array([[1.0, 0.0]])
    h = np.full(1, t / n_steps)
    for _ in range(n_steps):
        out = step(harmonic, s, h)
        s = out[0] if isinstance(out, tuple) else out
    return np.max(np.abs(s - np.array([[np.cos(t), -np.sin(t)]])))


def test_orders():
    e = [_run(orc.dormand_prince_step, n) for n in (50, 100, 200)]
    assert 4.6 < np.log2(e[0] / e[1]) < 5.4 and 4.6 < np.log2(e[1] / e[2]) < 5.4
    e = [_run(orc.rk4_step, n) for n in (100, 200, 400)]
    assert 3.7 < np.log2(e[0] / e[1]) < 4.3 and 3.7 < np.log2(e[1] / e[2]) < 4.3
    # implicit midpoint: second order, and with enough sweeps it is the exact
    # midpoint rule, which for a linear oscillator is the Cayley map
    # (norm-preserving: |y| stays 1 to rounding)
    mid = lambda f, s, h: orc.implicit_midpoint_step(f, s, h, iterations=30)  # noqa: E731
    e = [_run(mid, n) for n in (100, 200, 400)]
    assert 1.9 < np.log2(e[0] / e[1]) < 2.1 and 1.9 < np.log2(e[1] / e[2]) < 2.1
    s = np.array([[1.0, 0.0]])
    for _ in range(1000):
        s = mid(harmonic, s, np.full(1, 0.3))
    assert abs(np.hypot(s[0, 0], s[0, 1]) - 1.0) < 1e-12


def test_tableau_order_conditions():
    """Exact rational Dormand-Prince coefficients satisfy the order
    conditions up to 5 (b) and 4 (b - e); the doubles equal them rounded."""
    a = {(2, 1): F(1, 5), (3, 1): F(3, 40), (3, 2): F(9, 40), (4, 1): F(44, 45),
         (4, 2): F(-56, 15), (4, 3): F(32, 9), (5, 1): F(19372, 6561), (5, 2): F(-25360, 2187),
         (5, 3): F(64448, 6561), (5, 4): F(-212, 729), (6, 1): F(9017, 3168),
         (6, 2): F(-355, 33), (6, 3): F(46732, 5247), (6, 4): F(49, 176),
         (6, 5): F(-5103, 18656), (7, 1): F(35, 384), (7, 3): F(500, 1113),
         (7, 4): F(125, 192), (7, 5): F(-2187, 6784), (7, 6): F(11, 84)}
    c = {i: sum((v for (r, _), v in a.items() if r == i), F(0)) for i in range(2, 8)}
    c[1] = F(0)
    b = {j: a.get((7, j), F(0)) for j in range(1, 8)}
    e = {1: F(71, 57600), 2: F(0), 3: F(-71, 16695), 4: F(71, 1920), 5: F(-17253, 339200),
         6: F(22, 525), 7: F(-1, 40)}
    b4 = {j: b[j] - e[j] for j in range(1, 8)}
    for w, order in ((b, 5), (b4, 4)):
        assert sum(w.values()) == 1
        for k in range(1, order):  # bushy trees: sum w c^k = 1/(k+1)
            assert sum(w[j] * c[j] ** k for j in range(1, 8)) == F(1, k + 1)
    assert c[7] == 1 and c[2] == F(1, 5) and c[6] == 1
    for (i, j), v in a.items():
        name = {7: "B"}.get(i, f"A{i}")
        if i == 7:
            val = getattr(orc, f"B{j}")
        else:
            val = getattr(orc, f"A{i}{j}")
        assert val == float(v), (i, j, name)
    for j, v in e.items():
        if v:
            assert getattr(orc, f"E{j}") == float(v)


def test_error_estimate_band_and_nan():
    s = np.array([[1.0, 0.0]])
    new, err = orc.dormand_prince_step(harmonic, s, np.array([0.3]))
    true = np.max(np.abs(new - np.array([[np.cos(0.3), -np.sin(0.3)]])))
    assert 1e-2 < np.max(np.abs(err)) / max(true, 1e-300) < 1e2
    r = orc.error_ratio(s, s, np.array([[np.nan, 0.0]]), rtol=1e-9, atol=1e-12)
    assert np.isinf(r[0])


def test_error_ratio_and_step_factor_known_answers():
    y = np.array([[2.0, -4.0]])
    yn = np.array([[3.0, -1.0]])
    err = np.array([[3e-6, 2e-6]])
    r = orc.error_ratio(y, yn, err, rtol=1e-6, atol=0.0)
    # scales 3e-6 and 4e-6 -> sqrt((1 + 0.25)/2)
    np.testing.assert_allclose(r, [np.sqrt(0.625)], rtol=1e-15)
    f = orc.step_factor(np.array([1.0, 1e-30, np.inf, 32.0 ** 5 / 1e5]))
    np.testing.assert_allclose(f, [0.9, 5.0, 0.2, 0.9 * (32.0 ** 5 / 1e5) ** -0.2], rtol=1e-15)


def _row_sum(values):
    """The device kernels' row sum (error_ratio_kernel): sequential from 0
    below 8 values, else 8 interleaved partial sums joined pairwise, then the
    remainder sequentially."""
    d = len(values)
    if d < 8:
        s = 0.0
        for v in values:
            s += v
        return s
    r = [float(v) for v in values[:8]]
    k = 8
    while k < d - d % 8:
        for j in range(8):
            r[j] += values[k + j]
        k += 8
    s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
    for v in values[k:]:
        s += v
    return s


def test_row_sum_order_is_numpys():
    """NumPy's mean over the last axis (the reference's error_ratio,
    integrators.py:116) sums each row of up to 128 values in the order the
    device kernels use, bit for bit; at 129 values it splits differently,
    which is why the device entry takes at most 128."""
    rng = np.random.default_rng(3)
    for d in list(range(1, 34)) + [63, 64, 65, 100, 127, 128]:
        x = rng.random((64, d)) * np.exp(rng.normal(0.0, 5.0, (64, d)))
        mine = np.array([_row_sum([float(v) for v in row]) for row in x]) / float(d)
        np.testing.assert_array_equal(mine, np.mean(x, axis=-1), err_msg=f"d={d}")
    x = rng.random((64, 129)) * np.exp(rng.normal(0.0, 5.0, (64, 129)))
    mine = np.array([_row_sum([float(v) for v in row]) for row in x]) / 129.0
    assert (mine != np.mean(x, axis=-1)).any()
