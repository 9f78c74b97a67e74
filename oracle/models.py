"""Spin-1/2 model operators and imaginary-time gates, restating the math of
TeNeS's input generators (tool/tenes_simple.py:703-798 model_bondhamiltonian,
tool/tenes_std.py:893-908 make_evolution_twosite) for the square lattice
(coordination z = 4; bonds: every site's +x bond, then every site's +y bond,
tool/tenes_simple.py:256-257)."""
from __future__ import annotations

import numpy as np

Sz = np.array([[0.5, 0.0], [0.0, -0.5]])
Sx = np.array([[0.0, 0.5], [0.5, 0.0]])
Sy = np.array([[0.0, -0.5j], [0.5j, 0.0]])
Id2 = np.eye(2)


def bond_hamiltonian(Jz=0.0, Jx=0.0, Jy=0.0, hx=0.0, hy=0.0, hz=0.0, z=4):
    """ham[in1, in2, out1, out2] (tenes_simple.py:764-796), fields split over
    the z bonds of a site."""
    hx, hy, hz = hx / z, hy / z, hz / z
    ham = (Jz * np.einsum("ac,bd->abcd", Sz, Sz) + Jx * np.einsum("ac,bd->abcd", Sx, Sx)
           + Jy * np.real(np.einsum("ac,bd->abcd", Sy, Sy)))
    for h, S in ((hz, Sz), (hx, Sx), (hy, Sy)):
        if h != 0.0:
            ham = ham - h * (np.einsum("ac,bd->abcd", S, Id2) + np.einsum("ac,bd->abcd", Id2, S))
    return ham


def evolution_twosite(ham: np.ndarray, tau: float) -> np.ndarray:
    """tenes_std.py:893-908: exp(-tau H) through eigh of H reshaped to
    (d1 d2) x (d1 d2).  A complex tau (tau = i t) gives the real-time gate
    exp(-i t H) of TeNeS's time-evolution mode."""
    d = ham.shape
    H = ham.reshape(d[0] * d[1], d[0] * d[1])
    w, V = np.linalg.eigh(H)
    evo = (V * np.exp(-tau * w)) @ V.conj().T
    evo = evo.reshape(d)
    real = np.isrealobj(ham) and np.imag(tau) == 0.0
    return evo.real.copy() if real or np.max(np.abs(evo.imag)) == 0.0 else evo


def square_bonds(lat):
    """(source, source_leg) of every nearest-neighbour bond, +x first."""
    return [(s, 2) for s in range(lat.N_UNIT)] + [(s, 1) for s in range(lat.N_UNIT)]


def neel_dirs(lat):
    """Neel initial vectors on the (x + y) parity sublattices."""
    return [[1.0, 0.0] if (lat.x(i) + lat.y(i)) % 2 == 0 else [0.0, 1.0] for i in range(lat.N_UNIT)]
