"""Host-side input preparation for the GPU path (not on the measured path):
initial site tensors in TeNeS's RNG order (src/iTPS/tensors.cpp:83-128:
std::mt19937 + uniform_real_distribution(-1, 1), first leg fastest), spin-1/2
bond Hamiltonians and the exp(-tau H) gates of TeNeS's input generators
(tool/tenes_simple.py:764-796, tool/tenes_std.py:893-908).  The simple update
itself runs on the GPU (Environment.simple_update)."""
from __future__ import annotations

import numpy as np

SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
SX = np.array([[0.0, 0.5], [0.5, 0.0]])
SY = np.array([[0.0, -0.5j], [0.5j, 0.0]])


def _uniform(gen_state: np.random.MT19937, n: int) -> np.ndarray:
    """libstdc++ uniform_real_distribution<double>(-1, 1): two 32-bit draws
    per double, (x0 + x1 2^32) / 2^64 mapped to [-1, 1)."""
    r = gen_state.random_raw(2 * n).astype(np.uint64).reshape(n, 2)
    s = r[:, 0].astype(np.float64) + r[:, 1].astype(np.float64) * 4294967296.0
    c = s / 18446744073709551616.0
    c = np.where(c >= 1.0, np.nextafter(1.0, 0.0), c)
    return -1.0 + 2.0 * c


def _mt19937(seed: int) -> np.random.MT19937:
    g = np.random.MT19937(0)
    g._legacy_seeding(int(seed) & 0xFFFFFFFF)
    return g


def initial_tensors(n_sites: int, D: int, d: int, dirs, noise: float, seed: int = 11):
    """Real site tensors (D, D, D, D, d): entry (0,0,0,0,:) = dirs[i], the rest
    noise * U(-1, 1)."""
    gen = _mt19937(seed)
    out = []
    for i in range(n_sites):
        vals = noise * _uniform(gen, D ** 4 * d)
        T = vals.reshape(d, D, D, D, D).transpose(4, 3, 2, 1, 0).copy()
        T[0, 0, 0, 0, :] = dirs[i]
        out.append(T)
    return out


def random_tensors(n_sites: int, D: int, d: int = 2, seed: int = 11):
    """Real site tensors (D, D, D, D, d) with i.i.d. U(-1, 1) entries, drawn
    from one std::mt19937(seed) stream site by site in TeNeS's fill order
    (first leg fastest), each scaled to unit Frobenius norm."""
    gen = _mt19937(seed)
    out = []
    for _ in range(n_sites):
        T = _uniform(gen, D ** 4 * d).reshape(d, D, D, D, D).transpose(4, 3, 2, 1, 0).copy()
        out.append(T / np.linalg.norm(T))
    return out


def bond_hamiltonian(Jz=0.0, Jx=0.0, Jy=0.0, hx=0.0, hz=0.0, z=4):
    """H[in1, in2, out1, out2] = Jz SzSz + Jx SxSx + Jy Re(SySy)
    - (hx/z)(Sx 1 + 1 Sx) - (hz/z)(Sz 1 + 1 Sz)."""
    e = np.eye(2)
    k = lambda A, B: np.einsum("ac,bd->abcd", A, B)  # noqa: E731
    H = Jz * k(SZ, SZ) + Jx * k(SX, SX) + Jy * np.real(k(SY, SY))
    if hx:
        H = H - (hx / z) * (k(SX, e) + k(e, SX))
    if hz:
        H = H - (hz / z) * (k(SZ, e) + k(e, SZ))
    return H


def gate(H: np.ndarray, tau: float) -> np.ndarray:
    d = H.shape
    w, V = np.linalg.eigh(H.reshape(d[0] * d[1], -1))
    return ((V * np.exp(-tau * w)) @ V.conj().T).real.reshape(d)


def square_bonds(LX: int, LY: int):
    """(source, source_leg): every +x bond (leg 2), then every +y bond (leg 1)."""
    n = LX * LY
    return [(s, 2) for s in range(n)] + [(s, 1) for s in range(n)]


def neel_dirs(LX: int, LY: int):
    return [[1.0, 0.0] if ((i % LX) + (i // LX)) % 2 == 0 else [0.0, 1.0] for i in range(LX * LY)]
