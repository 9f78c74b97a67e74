"""Full update restated in numpy (TeNeS src/iTPS/core/full_update.cpp:85-441
and the driver src/iTPS/full_update.cpp:29-186, fast-full-update moves)."""
from __future__ import annotations

import numpy as np

from . import ctm
from .contract import Create_Environment_two_sites, trace

td = np.tensordot
cj = np.conj


class FUParameters:
    """Full-update fields of PEPS_Parameters (PEPS_Parameters.cpp:55-64)."""

    def __init__(self, **kw):
        self.Inverse_Env_cut = 1e-12
        self.Full_Inverse_precision = 1e-12
        self.Full_Convergence_Epsilon = 1e-6
        self.Full_max_iteration = 100
        self.Full_Gauge_Fix = True
        self.Full_Use_FastFullUpdate = True
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)


def _qr(A, row_legs, col_legs):
    B = A.transpose(list(row_legs) + list(col_legs))
    rs = [A.shape[i] for i in row_legs]
    cs = [A.shape[i] for i in col_legs]
    Q, R = np.linalg.qr(B.reshape(int(np.prod(rs)), int(np.prod(cs))), mode="reduced")
    k = Q.shape[1]
    return Q.reshape(rs + [k]), R.reshape([k] + cs)


def _svd(A, row_legs, col_legs):
    B = A.transpose(list(row_legs) + list(col_legs))
    rs = [A.shape[i] for i in row_legs]
    cs = [A.shape[i] for i in col_legs]
    U, s, Vh = np.linalg.svd(B.reshape(int(np.prod(rs)), int(np.prod(cs))), full_matrices=False)
    k = s.size
    return U.reshape(rs + [k]), s, Vh.reshape([k] + cs)


def _inv_vals(s, cut):
    s = s.copy()
    smax = s[0]
    return np.where(s / smax > cut, 1.0 / np.where(s > 0, s, 1.0), 0.0)


def _delta(Env, Theta, R1, R2):
    return (-2.0 * trace(td(td(Env, Theta, ([0, 1], [0, 1])), cj(R2), ([1, 3], [0, 2])), cj(R1),
                         [0, 1, 2], [0, 2, 1])
            + trace(R1, td(R2, td(Env, td(cj(R1), cj(R2), ([1], [1])), ([2, 3], [0, 2])), ([0, 2], [1, 3])),
                    [0, 1, 2], [1, 0, 2]))


def Full_update_bond_horizontal(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, Tn1, Tn2, op12,
                                p: FUParameters):
    """core/full_update.cpp:85-393"""
    D_connect = Tn1.shape[2]
    Q1, R1 = _qr(Tn1, (0, 1, 3), (2, 4))
    Q2, R2 = _qr(Tn2, (1, 2, 3), (0, 4))
    envR1, envR2 = R1.shape[0], R2.shape[0]
    Theta = td(td(R1, R2, ([1], [1])), op12, ([1, 3], [0, 1]))
    Env = Create_Environment_two_sites(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, Q1,
                                       Q2.transpose(3, 0, 1, 2))
    Env = 0.5 * (Env + cj(Env.transpose(2, 3, 0, 1)))
    sh = Env.shape
    w, Z = np.linalg.eigh(Env.reshape(sh[0] * sh[1], -1))
    Z = Z.reshape(sh[0], sh[1], -1)
    w_max = abs(w[envR1 * envR2 - 1])
    w = np.where(w / w_max > p.Inverse_Env_cut, np.sqrt(np.where(w > 0, w, 0.0)), 0.0)
    Z = Z * w
    if p.Full_Gauge_Fix:
        _, LR1 = _qr(Z, (2, 1), (0,))
        _, LR2 = _qr(Z, (2, 0), (1,))
        Theta = td(LR1, td(LR2, Theta, ([1], [1])), ([1], [1]))
        u, s, vt = np.linalg.svd(LR1, full_matrices=False)
        u = u * _inv_vals(s, p.Inverse_Env_cut)
        LR1_inv = td(cj(u), cj(vt), ([1], [0]))
        u, s, vt = np.linalg.svd(LR2, full_matrices=False)
        u = u * _inv_vals(s, p.Inverse_Env_cut)
        LR2_inv = td(cj(u), cj(vt), ([1], [0]))
        Z = td(td(Z, LR1_inv, ([0], [1])), LR2_inv, ([0], [1]))
        Env = td(Z, cj(Z), ([0], [0]))
    else:
        Env = td(Z, cj(Z), ([2], [2]))
    U, s, VT = _svd(Theta, (0, 2), (1, 3))
    lam = s[:D_connect]
    U = U[:, :, :D_connect]
    VT = VT[:D_connect]
    lam = np.sqrt(lam / np.sqrt(np.sum(lam * lam)))
    U = U * lam
    VT = VT * lam[:, None, None]
    R1 = U.transpose(0, 2, 1)
    R2 = VT.transpose(1, 0, 2)
    C_phi = trace(td(Env, Theta, ([0, 1], [0, 1])), cj(Theta), [0, 1, 2, 3], [0, 1, 2, 3])
    old = _delta(Env, Theta, R1, R2)
    count = 0
    conv = False
    while not conv and count < p.Full_max_iteration:
        W = td(td(Env, Theta, ([0, 1], [0, 1])), cj(R2), ([1, 3], [0, 2]))
        N = td(td(Env, R2, ([1], [0])), cj(R2), ([2, 4], [0, 2]))
        Uu, s, Vv = _svd(N, (1, 3), (0, 2))
        Uu = Uu * _inv_vals(s, p.Full_Inverse_precision)
        R1 = td(td(cj(Uu), cj(Vv), ([2], [0])), W, ([0, 1], [0, 2]))
        W = td(td(Env, Theta, ([0, 1], [0, 1])), cj(R1), ([0, 2], [0, 2]))
        N = td(td(Env, R1, ([0], [0])), cj(R1), ([1, 4], [0, 2]))
        Uu, s, Vv = _svd(N, (1, 3), (0, 2))
        Uu = Uu * _inv_vals(s, p.Full_Inverse_precision)
        R2 = td(td(cj(Uu), cj(Vv), ([2], [0])), W, ([0, 1], [0, 2]))
        delta = _delta(Env, Theta, R1, R2)
        if abs(old - delta) / abs(C_phi) < p.Full_Convergence_Epsilon:
            conv = True
        old = delta
        count += 1
    if p.Full_Gauge_Fix:
        R1 = td(LR1_inv, R1, ([0], [0]))
        R2 = td(LR2_inv, R2, ([0], [0]))
    q1, r1 = _qr(R1, (0, 2), (1,))
    q2, r2 = _qr(R2, (0, 2), (1,))
    U, s, VT = np.linalg.svd(td(r1, r2, ([1], [1])), full_matrices=False)
    s = s.copy()
    nrm = np.sqrt(np.sum(s[:D_connect] ** 2))
    s[:D_connect] = np.sqrt(s[:D_connect] / nrm)
    U = U * s
    VT = VT * s[:, None]
    R1 = td(q1, U, ([2], [0]))
    R2 = td(q2, VT, ([2], [1]))
    T1 = td(Q1, R1, ([3], [0])).transpose(0, 1, 4, 2, 3)
    T2 = td(Q2, R2, ([3], [0])).transpose(4, 0, 1, 2, 3)
    return T1, T2


_ROT = {0: (2, 3, 0, 1, 4), 1: (3, 0, 1, 2, 4), 2: None, 3: (1, 2, 3, 0, 4)}
_BACK = {0: (2, 3, 0, 1, 4), 1: (1, 2, 3, 0, 4), 2: None, 3: (3, 0, 1, 2, 4)}


def Full_update_bond(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, Tn1, Tn2, op12, connect1, p):
    """core/full_update.cpp:395-441"""
    r = _ROT[connect1]
    T1 = Tn1 if r is None else Tn1.transpose(r)
    T2 = Tn2 if r is None else Tn2.transpose(r)
    n1, n2 = Full_update_bond_horizontal(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, T1, T2, op12, p)
    b = _BACK[connect1]
    if b is not None:
        n1, n2 = n1.transpose(b), n2.transpose(b)
    return n1, n2


def full_update_step(env, Tn, lat, source, leg, op, p_ctm, p_fu):
    """src/iTPS/full_update.cpp:29-152 for one two-site evolution operator."""
    target = lat.neighbor(source, leg)
    e = env
    if leg == 0:
        args = (e.C4[source], e.C2[target], e.C1[target], e.C3[source], e.eTb[source], e.eTb[target],
                e.eTl[target], e.eTt[target], e.eTt[source], e.eTr[source])
    elif leg == 1:
        args = (e.C4[source], e.C1[target], e.C2[target], e.C3[source], e.eTl[source], e.eTl[target],
                e.eTt[target], e.eTr[target], e.eTr[source], e.eTb[source])
    elif leg == 2:
        args = (e.C1[source], e.C2[target], e.C3[target], e.C4[source], e.eTt[source], e.eTt[target],
                e.eTr[target], e.eTb[target], e.eTb[source], e.eTl[source])
    else:
        args = (e.C2[source], e.C3[target], e.C4[target], e.C1[source], e.eTr[source], e.eTr[target],
                e.eTb[target], e.eTl[target], e.eTl[source], e.eTt[source])
    n1, n2 = Full_update_bond(*args, Tn[source], Tn[target], op, leg, p_fu)
    Tn[source] = n1
    Tn[target] = n2
    LX = lat.LX
    if p_fu.Full_Use_FastFullUpdate:
        sx, tx = source % LX, target % LX
        sy, ty = source // LX, target // LX
        if leg == 0:
            ctm.Right_move(env, Tn, sx, p_ctm, lat)
            ctm.Left_move(env, Tn, tx, p_ctm, lat)
        elif leg == 1:
            ctm.Bottom_move(env, Tn, sy, p_ctm, lat)
            ctm.Top_move(env, Tn, ty, p_ctm, lat)
        elif leg == 2:
            ctm.Left_move(env, Tn, sx, p_ctm, lat)
            ctm.Right_move(env, Tn, tx, p_ctm, lat)
        else:
            ctm.Top_move(env, Tn, sy, p_ctm, lat)
            ctm.Bottom_move(env, Tn, ty, p_ctm, lat)
        return env
    env, _, _ = ctm.Calc_CTM_Environment(Tn, p_ctm, lat)
    return env


def full_update(Tn, lat, ops, nsteps, p_ctm, p_fu):
    """src/iTPS/full_update.cpp:154-186: CTM from scratch, then nsteps passes
    over the operators (fast full update moves in between)."""
    if nsteps <= 0:
        return Tn
    env, _, _ = ctm.Calc_CTM_Environment(Tn, p_ctm, lat)
    for _ in range(nsteps):
        for source, leg, op in ops:
            env = full_update_step(env, Tn, lat, source, leg, op, p_ctm, p_fu)
    return Tn
