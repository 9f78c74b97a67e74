"""Simple update restated in numpy (TeNeS src/iTPS/core/simple_update.cpp:50-198,
src/iTPS/simple_update.cpp:25-138) and the initial-tensor construction of
src/iTPS/tensors.cpp:38-137."""
from __future__ import annotations

import numpy as np

from .rng import MT19937, UniformReal

td = np.tensordot

# leg order moved to the end before the QR, and the transpose that puts the
# new bond back (core/simple_update.cpp:97-120 / 166-197)
_TO_END = {0: (1, 2, 3, 0, 4), 1: (0, 2, 3, 1, 4), 2: (0, 1, 3, 2, 4), 3: (0, 1, 2, 3, 4)}
_BACK1 = {0: (4, 0, 1, 2, 3), 1: (0, 4, 1, 2, 3), 2: (0, 1, 4, 2, 3), 3: (0, 1, 2, 4, 3)}
_BACK2 = {0: (3, 0, 1, 2, 4), 1: (0, 3, 1, 2, 4), 2: (0, 1, 3, 2, 4), 3: (0, 1, 2, 3, 4)}


def _qr(A: np.ndarray, nrow_legs: int):
    """mptensor qr(A, rows, cols, Q, R): thin QR of the row/col matricisation."""
    shp = A.shape
    m = int(np.prod(shp[:nrow_legs]))
    Q, R = np.linalg.qr(A.reshape(m, -1), mode="reduced")
    k = Q.shape[1]
    return Q.reshape(shp[:nrow_legs] + (k,)), R.reshape((k,) + shp[nrow_legs:])


def _mulvec(T, vecs_by_leg):
    for leg, v in vecs_by_leg:
        shape = [1] * T.ndim
        shape[leg] = len(v)
        T = T * np.asarray(v).reshape(shape)
    return T


def Simple_update_bond(Tn1, Tn2, lambda1, lambda2, op12, connect1, inverse_lambda_cut=1e-12):
    """core/simple_update.cpp:50-198.  Returns (Tn1_new, Tn2_new, lambda_c)."""
    connect2 = (connect1 + 2) % 4

    def inv(lams):
        return [np.where(np.asarray(l) > inverse_lambda_cut, 1.0 / np.maximum(np.asarray(l), 1e-300), 0.0)
                for l in lams]

    l1inv = inv(lambda1)
    l2inv = inv(lambda2)
    dc = Tn1.shape[connect1]
    others1 = [i for i in range(4) if i != connect1]
    others2 = [i for i in range(4) if i != connect2]
    T1 = _mulvec(Tn1, [(i, lambda1[i]) for i in others1]).transpose(_TO_END[connect1])
    T2 = _mulvec(Tn2, [(i, lambda2[i]) for i in others2]).transpose(_TO_END[connect2])
    Q1, R1 = _qr(T1, 3)
    Q2, R2 = _qr(T2, 3)
    Theta = td(td(R1, R2, ([1], [1])), op12, ([1, 3], [0, 1]))
    th = Theta.shape
    U, s, VT = np.linalg.svd(Theta.transpose(0, 2, 1, 3).reshape(th[0] * th[2], th[1] * th[3]),
                             full_matrices=False)
    U = U.reshape(th[0], th[2], -1)
    VT = VT.reshape(-1, th[1], th[3])
    lam = s[:dc]
    Uc = U[:, :, :dc]
    VTc = VT[:dc]
    norm = np.sqrt(np.sum(lam * lam))
    lam = np.sqrt(lam / norm)
    Uc = Uc * lam
    VTc = VTc * lam[:, None, None]
    Q1 = _mulvec(Q1, [(k, l1inv[i]) for k, i in enumerate(others1)])
    Q2 = _mulvec(Q2, [(k, l2inv[i]) for k, i in enumerate(others2)])
    T1n = td(Q1, Uc, ([3], [0])).transpose(_BACK1[connect1])
    T2n = td(Q2, VTc, ([3], [1])).transpose(_BACK2[connect2])
    return T1n, T2n, lam


def initial_tensors(lat, pdims, vdims, initial_dirs, noises, seed=11, is_real=True):
    """tensors.cpp:83-128: Tn[i][0,0,0,0,:] = dir (random if all zero),
    other entries noise * U(-1,1) drawn first-leg-fastest from mt19937(seed)
    (imaginary parts from a second generator seeded seed*11+137)."""
    gen = MT19937(seed)
    gen_im = MT19937(seed * 11 + 137)
    dist = UniformReal(-1.0, 1.0)
    Tn = []
    for i in range(lat.N_UNIT):
        pd = pdims[i]
        vd = vdims[i]
        ndim = vd[0] * vd[1] * vd[2] * vd[3] * pd
        # ran_re[j] and ran_im[j] come from two independent generators
        re = dist(gen, ndim)
        im = dist(gen_im, ndim)
        d = list(initial_dirs[i])
        dim_ = np.zeros(pd)
        if all(x == 0.0 for x in d):
            d = [0.0] * pd
            for j in range(pd):
                d[j] = dist(gen)[0]
                dim_[j] = dist(gen_im)[0]
        # first-leg-fastest linear index nr -> tensor index
        vals = noises[i] * (re + 1j * im)
        T = vals.reshape(pd, vd[3], vd[2], vd[1], vd[0]).transpose(4, 3, 2, 1, 0).copy()
        T[0, 0, 0, 0, :] = np.asarray(d) + 1j * dim_
        Tn.append(T.real.copy() if is_real else T)
    return Tn


def simple_update_steps(Tn, lambdas, lat, evolution_ops, nsteps, inverse_lambda_cut=1e-12):
    """simple_update.cpp:103-138 (group 0, no gauge fixing): apply the list of
    two-site evolution operators [(source, source_leg, op), ...] nsteps times."""
    for _ in range(nsteps):
        for source, leg, op in evolution_ops:
            target = lat.neighbor(source, leg)
            tleg = (leg + 2) % 4
            t1, t2, lam = Simple_update_bond(Tn[source], Tn[target], lambdas[source], lambdas[target],
                                             op, leg, inverse_lambda_cut)
            lambdas[source][leg] = lam
            lambdas[target][tleg] = lam
            Tn[source] = t1
            Tn[target] = t2
    return Tn, lambdas
