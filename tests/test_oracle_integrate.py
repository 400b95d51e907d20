"""Pins the oracle's integrate_dense to the reference's (golden vectors of
integrate_dense run on one ray at a time) and to the conservation laws of
the reference's test_orbits.py."""

import numpy as np

from conftest import golden
from oracle import bh_oracle as orc


def _cases():
    g = golden("dense_orbits.npz")
    k = 0
    while f"c{k}_state0" in g.files:
        yield k, g
        k += 1


def test_integrate_dense_matches_reference_per_ray():
    st = orc.Kerr(spin=0.9)
    for k, g in _cases():
        states = orc.integrate_dense(lambda y: orc.geodesic_rhs(st, y), g[f"c{k}_state0"],
                                     float(g[f"c{k}_lambda"]), rtol=float(g[f"c{k}_rtol"]))
        assert len(states) - 1 == int(g[f"c{k}_steps"])
        np.testing.assert_array_equal(states[-1], g[f"c{k}_final"])


def test_golden_finals_conserve_invariants():
    st = orc.Kerr(spin=0.9)
    for k, g in _cases():
        s0, s1 = g[f"c{k}_state0"], g[f"c{k}_final"]
        e0, lz0 = orc.conserved_quantities(s0)
        e1, lz1 = orc.conserved_quantities(s1)
        np.testing.assert_array_equal(e1, e0)  # dp_t / dlambda = 0 exactly
        np.testing.assert_allclose(lz1, lz0, rtol=1e-7)
        h0 = orc.hamiltonian(st, s0)
        np.testing.assert_allclose(orc.hamiltonian(st, s1), h0, atol=1e-8)
    # circular orbit at r = 6 M stays at r = 6 M (test_orbits.py:108-136)
    g = golden("dense_orbits.npz")
    s1 = g["c1_final"]
    np.testing.assert_allclose(st.kerr_schild_radius(s1[:, 1], s1[:, 2], s1[:, 3]), 6.0,
                               rtol=1e-6)
    assert abs(s1[0, 3]) < 1e-8


def test_integrate_dense_matches_reference_on_workload_orbits():
    """Two of the timelike-orbit workload's orbits (c5_orbits.npz: proper
    time 1e4 M, rtol 1e-10), bit for bit; all eight golden orbits conserve
    E exactly, L_z to 1e-9 and H to 1e-10 over 5,000-7,000 steps."""
    st = orc.Kerr(spin=0.9)
    g = golden("c5_orbits.npz")
    for k in (0, 3):
        states = orc.integrate_dense(lambda y: orc.geodesic_rhs(st, y), g[f"o{k}_state0"],
                                     1.0e4, rtol=1e-10)
        assert len(states) - 1 == int(g[f"o{k}_steps"])
        np.testing.assert_array_equal(states[-1], g[f"o{k}_final"])
    k = 0
    while f"o{k}_index" in g.files:
        s0, s1 = g[f"o{k}_state0"], g[f"o{k}_final"]
        e0, lz0 = orc.conserved_quantities(s0)
        e1, lz1 = orc.conserved_quantities(s1)
        np.testing.assert_array_equal(e1, e0)
        np.testing.assert_allclose(lz1, lz0, rtol=1e-9)
        np.testing.assert_allclose(orc.hamiltonian(st, s1), orc.hamiltonian(st, s0), atol=1e-10)
        k += 1
    assert k == 8
