"""The oracle of the bond energy and its gradient (oracle/adjoint.py), pinned
on the CPU: its bond values against the contraction trees of oracle/contract.py
(which the reference's golden outputs pin), its gradient against central
finite differences, and both against the closed form of an embedded product
state (every observable of a product state is a Rayleigh quotient)."""
from __future__ import annotations

import numpy as np

from ctm_cases import random_tensors
from oracle import adjoint as A
from oracle import ctm, models
from oracle.itps import measure_twosite
from oracle.lattice import SquareLattice

LAT = SquareLattice(2, 2)
NN = ((1, 0), (0, 1), (-1, 0), (0, -1))


def _env(Tn, chi, sweeps):
    p = ctm.PEPSParameters(CHI=chi, Max_CTM_Iteration=sweeps, CTM_Convergence_Epsilon=0.0)
    env, _, _ = ctm.Calc_CTM_Environment(list(Tn), p, LAT)
    return env


def _heisenberg_bonds():
    H = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    return [(s, 1, 0, H) for s in range(4)] + [(s, 0, 1, H) for s in range(4)]


def test_bond_values_match_contraction_trees():
    """Normalised bond values of a non-symmetric operator, every source and
    direction, against measure_twosite (op12 tree / 1x2 or 2x1 norm)."""
    Tn = random_tensors(2, 2, seed=3)
    env = _env(Tn, 6, 3)
    op = np.random.default_rng(0).standard_normal((2, 2, 2, 2))
    eye = np.einsum("ac,bd->abcd", np.eye(2), np.eye(2))
    worst = 0.0
    for src in range(4):
        for dx, dy in NN:
            v = A.bond_value(env, Tn, LAT, src, dx, dy, op) / A.bond_value(env, Tn, LAT, src, dx, dy, eye)
            ref, _ = measure_twosite(env, Tn, LAT, [{"name": "x", "group": 0, "source": src, "dx": dx,
                                                      "dy": dy, "op": op}], [])
            worst = max(worst, abs(v - ref[0][(src, dx, dy)].real))
    assert worst < 1e-13, worst


def test_gradient_central_differences():
    """Directional derivatives along random directions against central
    differences (h = 1e-5, truncation error ~h^2), environment fixed."""
    Tn = random_tensors(2, 2, seed=5)
    env = _env(Tn, 6, 3)
    bonds = _heisenberg_bonds()
    _, g = A.energy_and_grad(env, list(Tn), LAT, bonds)
    rng = np.random.default_rng(1)
    h = 1e-5
    for _ in range(4):
        V = [rng.standard_normal(T.shape) for T in Tn]
        Ep, _ = A.energy_and_grad(env, [T + h * v for T, v in zip(Tn, V)], LAT, bonds, want_grad=False)
        Em, _ = A.energy_and_grad(env, [T - h * v for T, v in zip(Tn, V)], LAT, bonds, want_grad=False)
        fd = (Ep - Em) / (2 * h)
        an = sum(float(np.sum(x * v)) for x, v in zip(g, V))
        assert abs(fd - an) < 1e-6 * abs(an), (fd, an)


def test_product_state_closed_form():
    """Tn nonzero only at virtual index (0,0,0,0) with physical vectors v_i:
    the environment has rank 1, E = (1/4) sum_b e_b with the Rayleigh
    quotient e_b = <v1 v2|H|v1 v2> / (|v1|^2 |v2|^2), the gradient on the
    physical components is that of the quotients and zero elsewhere."""
    D = 2
    rng = np.random.default_rng(7)
    vs = [rng.standard_normal(2) for _ in range(4)]
    Tn = []
    for v in vs:
        T = np.zeros((D, D, D, D, 2))
        T[0, 0, 0, 0, :] = v
        Tn.append(T)
    env = _env(Tn, 4, 3)
    bonds = _heisenberg_bonds()
    E, g = A.energy_and_grad(env, Tn, LAT, bonds)
    H = bonds[0][3]
    E_ref = 0.0
    g_ref = [np.zeros(2) for _ in range(4)]
    for s, dx, dy, _ in bonds:
        t = LAT.other(s, dx, dy)
        v1, v2 = vs[s], vs[t]
        n = (v1 @ v1) * (v2 @ v2)
        e = np.einsum("ptqu,p,t,q,u->", H, v1, v2, v1, v2) / n
        E_ref += e
        d1 = (np.einsum("ktqu,t,q,u->k", H, v2, v1, v2) + np.einsum("ptku,p,t,u->k", H, v1, v2, v2)) / n
        d2 = (np.einsum("pkqu,p,q,u->k", H, v1, v1, v2) + np.einsum("ptqk,p,t,q->k", H, v1, v2, v1)) / n
        g_ref[s] += d1 - 2 * e * v1 / (v1 @ v1)
        g_ref[t] += d2 - 2 * e * v2 / (v2 @ v2)
    assert abs(E - E_ref / 4) < 1e-13
    for i in range(4):
        assert np.allclose(g[i][0, 0, 0, 0, :], g_ref[i] / 4, rtol=0, atol=1e-12)
        rest = g[i].copy()
        rest[0, 0, 0, 0, :] = 0
        assert np.max(np.abs(rest)) < 1e-12
