"""Pin the spin-1/2 bond Hamiltonians and gates of oracle.models and
tenes_amd.inputs (the inputs of every CTM parity test) to data TeNeS's own
test inputs hold, and the gate to the matrix exponential.

TeNeS's input generators (tool/tenes_simple.py:703-798 model_bondhamiltonian,
tool/tenes_std.py:893-908 make_evolution_twosite) wrote the operator elements
of test/data/*.toml, which tests/golden/make_fixtures.py copied as data:
  * FT_TFI_square: transverse-field Ising, Jz = -1, hx = 1 (bond Hamiltonian,
    and its imaginary-time gates at tau = 0.005);
  * TE_TFI: Jz = -1, hx = 0.5 (bond Hamiltonian, real-time gates exp(-i t H)
    at t = 0.01);
  * AntiferroHeisenberg_real / J1J2_AFH: Heisenberg J = 1 (and J2 = 0.1 on
    the diagonal bonds of J1J2), gates at tau = 0.01.
The transverse-field term is split over the z = 4 bonds of a square-lattice
site; a wrong z, sign, spin normalisation or tau fails these comparisons.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import golden_io as G
from oracle import models
from tenes_amd import inputs


def _op(fx, k, key="twosite"):
    o = fx[key][k]
    return G.tensor_from_elements([2, 2, 2, 2], o["elements"], False)


def _gate(fx, k, key="evolution"):
    o = fx[key][k]
    return G.tensor_from_elements(o["dimensions"], o["elements"], False)


def test_tfi_bond_hamiltonian_matches_ft_tfi_square_and_te_tfi():
    for name, hx in (("FT_TFI_square", 1.0), ("TE_TFI", 0.5)):
        ref = _op(G.load(name), 0)
        for H in (models.bond_hamiltonian(Jz=-1.0, hx=hx), inputs.bond_hamiltonian(Jz=-1.0, hx=hx)):
            assert np.max(np.abs(H - ref)) < 1e-15, name
        # a field split over z = 2 bonds, or of the opposite sign, is not TeNeS's
        assert np.max(np.abs(models.bond_hamiltonian(Jz=-1.0, hx=hx, z=2) - ref)) > 1e-2
        assert np.max(np.abs(models.bond_hamiltonian(Jz=-1.0, hx=-hx) - ref)) > 1e-2


def test_heisenberg_bond_hamiltonians_match_afh_and_j1j2():
    afh = G.load("AntiferroHeisenberg_real")
    j12 = G.load("J1J2_AFH")
    H1 = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    assert np.max(np.abs(H1 - _op(afh, 0))) < 1e-15
    assert np.max(np.abs(inputs.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0) - _op(afh, 0))) < 1e-15
    assert np.max(np.abs(H1 - _op(j12, 0))) < 1e-15
    assert np.max(np.abs(models.bond_hamiltonian(Jz=0.1, Jx=0.1, Jy=0.1) - _op(j12, 1))) < 1e-15
    # the SySy observable of the fixtures is Re(Sy (x) Sy)
    assert np.max(np.abs(models.bond_hamiltonian(Jy=1.0) - _op(afh, 3))) < 1e-15


def test_imaginary_time_gates_match_fixtures():
    ft = G.load("FT_TFI_square")
    g = models.evolution_twosite(models.bond_hamiltonian(Jz=-1.0, hx=1.0), ft["parameter"]["simple_update"]["tau"])
    for k in range(len(ft["evolution"])):
        assert np.max(np.abs(g - _gate(ft, k))) < 1e-15
    gi = inputs.gate(inputs.bond_hamiltonian(Jz=-1.0, hx=1.0), ft["parameter"]["simple_update"]["tau"])
    assert np.max(np.abs(gi - _gate(ft, 0))) < 1e-15
    afh = G.load("AntiferroHeisenberg_real")
    tau = afh["parameter"]["simple_update"]["tau"]
    g = models.evolution_twosite(models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0), tau)
    assert g[0, 0, 0, 0] == pytest.approx(np.exp(-0.0025), abs=1e-16)  # |up up> has energy 1/4
    for key in ("evolution", "evolution_full"):
        for k in range(len(afh[key])):
            assert np.max(np.abs(g - _gate(afh, k, key))) < 1e-15, (key, k)
    assert np.max(np.abs(inputs.gate(inputs.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0), tau) - _gate(afh, 0))) < 1e-15
    # the J1 gates of J1J2 are the same Heisenberg gates
    j12 = G.load("J1J2_AFH")
    for k in range(8):
        assert np.max(np.abs(g - _gate(j12, k))) < 1e-15


def test_real_time_gates_match_te_tfi():
    te = G.load("TE_TFI")
    t = te["parameter"]["simple_update"]["tau"]
    g = models.evolution_twosite(models.bond_hamiltonian(Jz=-1.0, hx=0.5), 1j * t)
    assert np.iscomplexobj(g)
    for k in range(len(te["evolution"])):
        assert np.max(np.abs(g - _gate(te, k))) < 1e-15


def test_j1j2_path_gates_compose_to_the_second_neighbour_gate():
    """J1J2's second-neighbour gates are path-decomposed (tenes_std.py): a
    (2,2,2,8) gate on (i, j) and an (8,2,2,2) gate on (j, k), the 8 = (j's
    physical leg, 4 auxiliary) carrying the operator's SVD index through the
    intermediate site.  Composed over that index each pair is
    exp(-tau J2 S_i.S_k) on (i, k) times the identity on j."""
    j12 = G.load("J1J2_AFH")
    tau = j12["parameter"]["simple_update"]["tau"]
    U = sl.expm(-tau * models.bond_hamiltonian(Jz=0.1, Jx=0.1, Jy=0.1).reshape(4, 4)).reshape(2, 2, 2, 2)
    ref = np.einsum("acpr,bq->abcpqr", U, np.eye(2))
    for key in ("evolution", "evolution_full"):
        ev = G.evolution(j12, True, key)
        assert len(ev) == 24
        for k in range(8, 24, 2):
            (_, _, g1), (_, _, g2) = ev[k], ev[k + 1]
            assert g1.shape == (2, 2, 2, 8) and g2.shape == (8, 2, 2, 2)
            T = np.einsum("abpX,Xcqr->abcpqr", g1, g2)
            assert np.max(np.abs(T - ref)) < 1e-14, (key, k)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gate_is_the_matrix_exponential(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((4, 4))
    Hr = (A + A.T).reshape(2, 2, 2, 2)
    B = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    Hc = (B + B.conj().T).reshape(2, 2, 2, 2)
    for H in (Hr, Hc):
        for tau in (0.01, 0.3, 0.05j):
            ref = sl.expm(-tau * H.reshape(4, 4)).reshape(2, 2, 2, 2)
            assert np.max(np.abs(models.evolution_twosite(H, tau) - ref)) < 1e-13
    assert np.max(np.abs(inputs.gate(Hr, 0.3) - sl.expm(-0.3 * Hr.reshape(4, 4)).reshape(2, 2, 2, 2))) < 1e-13
