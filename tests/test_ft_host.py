"""Host logic of the finite-temperature driver (tenes_amd/finite_temperature.py),
no GPU: the lattice arithmetic against the oracle's SquareLattice
(src/SquareLattice.hpp), the TPO initial state (tensors.cpp:154-203), the
energy density (density.cpp:39-76) against the oracle's, and the split of
complex operators into real parts."""
import numpy as np
import pytest

from oracle import itps
from oracle.lattice import SquareLattice
from tenes_amd import finite_temperature as FTG


@pytest.mark.parametrize("LX,LY,skew", [(2, 2, 0), (2, 1, 1), (3, 2, 1), (4, 2, 3), (3, 3, 2)])
def test_lattice_helpers_match_square_lattice(LX, LY, skew):
    lat = SquareLattice(LX, LY, skew)
    assert FTG._noskew(LX, LY, skew) == (lat.LX_noskew, lat.LY_noskew)
    for i in range(LX * LY):
        for dx in range(-5, 6):
            for dy in range(-5, 6):
                assert FTG._other(LX, LY, skew, i, dx, dy) == lat.other(i, dx, dy), (i, dx, dy)


def test_initial_tpo_is_identity_on_bond_zero():
    Tn = FTG.initial_tpo([2, 3], [[3, 1, 2, 2], [2, 2, 1, 4]])
    assert Tn[0].shape == (3, 1, 2, 2, 2, 2) and Tn[1].shape == (2, 2, 1, 4, 3, 3)
    for T in Tn:
        d = T.shape[4]
        assert np.array_equal(T[0, 0, 0, 0], np.eye(d))
        assert np.count_nonzero(T) == d


def test_density_matches_oracle():
    rng = np.random.default_rng(0)
    lat = SquareLattice(2, 2)
    pdims = [2, 2, 1, 2]  # a site of dimension 1 does not count
    one = {0: {i: rng.standard_normal() for i in range(4)}, 1: {i: rng.standard_normal() for i in range(4)}}
    two = {0: {(i, 1, 0): rng.standard_normal() for i in range(4)}, 1: {(0, 0, 1): rng.standard_normal()}}
    onames = {0: "site_hamiltonian", 1: "Sz"}
    tnames = {0: "bond_hamiltonian", 1: "SzSz"}
    got = FTG.density(one, two, pdims, onames, tnames)
    ref = itps.density(one, two, lat, pdims, onames, tnames)
    assert got.keys() == ref.keys()
    for k in ref:
        assert abs(got[k] - ref[k]) < 1e-15, k


def test_complex_operator_parts():
    Sy = np.array([[0.0, -0.5j], [0.5j, 0.0]])
    parts = FTG._parts(Sy)
    assert len(parts) == 2
    rebuilt = sum(c * p for p, c in parts)
    assert np.array_equal(rebuilt, Sy)
    assert len(FTG._parts(np.eye(2))) == 1
    assert len(FTG._parts(np.eye(2, dtype=complex))) == 1
