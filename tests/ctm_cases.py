"""Seeded inputs for the CTM parity tests (shared by CPU and GPU tests):
simple-update ground-state-like tensors (Heisenberg / transverse-field
Ising on the 2x2 square cell) and random unit-norm tensors."""
from __future__ import annotations

import functools

import numpy as np

from oracle import models, simple_update as SU
from oracle.lattice import SquareLattice


@functools.lru_cache(maxsize=None)
def su_tensors(model: str, D: int, steps: int, tau: float = 0.01, seed: int = 11, noise: float = 0.01,
               hx: float = 1.0):
    """Site tensors after `steps` simple-update steps on the 2x2 cell.
    model 'heisenberg': S.S from the Neel state; 'tfi': -SzSz - hx Sx from
    the ferro state."""
    lat = SquareLattice(2, 2)
    if model == "heisenberg":
        ham = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
        dirs = models.neel_dirs(lat)
    elif model == "tfi":
        ham = models.bond_hamiltonian(Jz=-1.0, hx=hx)
        dirs = [[1.0, 0.0]] * 4
    else:
        raise ValueError(model)
    gate = models.evolution_twosite(ham, tau)
    ops = [(s, leg, gate) for s, leg in models.square_bonds(lat)]
    Tn = SU.initial_tensors(lat, [2] * 4, [[D] * 4] * 4, dirs, [noise] * 4, seed=seed, is_real=True)
    lam = [[np.ones(D) for _ in range(4)] for _ in range(4)]
    Tn, lam = SU.simple_update_steps(Tn, lam, lat, ops, steps)
    return tuple(t.copy() for t in Tn), ham


def random_tensors(D: int, d: int = 2, seed: int = 7, n: int = 4):
    """Entries i.i.d. uniform[-1, 1], each tensor scaled to unit Frobenius norm."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        T = rng.uniform(-1.0, 1.0, size=(D, D, D, D, d))
        out.append(T / np.linalg.norm(T))
    return out


def spin_ops():
    return {"Sz": models.Sz, "Sx": models.Sx}


def spectra_diff(a, b):
    """Largest difference of two descending spectra, the shorter padded with
    zeros (with the projector cut on the device the GPU environment keeps CHI
    columns where the reference's shrinks: the extra singular values are 0)."""
    a, b = np.asarray(a), np.asarray(b)
    n = max(a.size, b.size)
    return float(np.max(np.abs(np.pad(a, (0, n - a.size)) - np.pad(b, (0, n - b.size)))))
