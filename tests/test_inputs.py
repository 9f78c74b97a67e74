"""Host-side input preparation of the package against the oracle."""
import numpy as np

from oracle import models, simple_update as SU
from oracle.lattice import SquareLattice
from tenes_amd import inputs


def test_initial_tensors_match_tenes_rng_order():
    lat = SquareLattice(2, 2)
    for D in (2, 5):
        ref = SU.initial_tensors(lat, [2] * 4, [[D] * 4] * 4, models.neel_dirs(lat), [0.01] * 4, seed=11)
        got = inputs.initial_tensors(4, D, 2, inputs.neel_dirs(2, 2), 0.01, seed=11)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)


def test_heisenberg_bond_spectrum():
    """Singlet energy -3/4, triplet 1/4 (the elements themselves are pinned to
    TeNeS's inputs in test_oracle_models.py)."""
    w = np.linalg.eigvalsh(inputs.bond_hamiltonian(Jz=1, Jx=1, Jy=1).reshape(4, 4))
    assert np.allclose(w, [-0.75, 0.25, 0.25, 0.25])
    w = np.linalg.eigvalsh(models.bond_hamiltonian(Jz=1, Jx=1, Jy=1).reshape(4, 4))
    assert np.allclose(w, [-0.75, 0.25, 0.25, 0.25])
