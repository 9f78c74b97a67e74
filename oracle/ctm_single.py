"""Single-layer CTMRG of an iTPO (density-operator tensor product), restated
in numpy from TeNeS src/iTPS/core/ctm_single.cpp: the environment of the
network whose site tensors are the TPO traced over its two physical legs.
Test infrastructure: the checker of the GPU library, never part of it.

Leg conventions:
  TPO site tensor (left, up, right, down, phys_ket, phys_bra);
  traced site tensor a = sum_p T[..., p, p]: (left, up, right, down);
  C (chi, chi); eT (chi, chi, D), with the same leg order as the iTPS
  edges of ctm.cpp:984-1192 minus the bra leg.
The sweep and the convergence check are the iTPS ones (ctm.cpp:815-969);
the projector SVD is LAPACK's (the RSVD branch, like oracle.ctm's, uses the
exact SVD).
"""
from __future__ import annotations

import numpy as np

from .ctm import Env, PEPSParameters, _svd, _truncate, max_abs, Check_Convergence_CTM

td = np.tensordot
ein = np.einsum


def Make_single_tensor_density(Tn):
    """ctm_single.cpp:43-50: trace of each TPO over its physical legs."""
    return [ein("lurdpp->lurd", T) for T in Tn]


def Calc_projector_left_block_single(C1, C4, eT1, eT6, eT7, eT8, T1, T4, p: PEPSParameters):
    """ctm_single.cpp:125-265 (full-SVD branch; t41 == 1 gives identity
    projectors).  Returns PU, PL with legs (e78, t41, cut)."""
    e12, e56, e78 = eT1.shape[1], eT6.shape[0], eT8.shape[0]
    t12, t41, t34 = T1.shape[2], T1.shape[3], T4.shape[2]
    if t41 == 1:
        I = np.eye(e78, dtype=C1.dtype)
        return I.reshape(e78, 1, e78), I.reshape(e78, 1, e78)
    # upper-left block, legs (e78, t41 | e12, t12)
    LT = ein("ab,bxu,yal,lurd->ydxr", C1, eT1, eT8, T1).reshape(e78 * t41, e12 * t12)
    # lower-left block, legs (e78, t41 | e56, t34)
    LB = ein("cf,fyl,zcd,lurd->yuzr", C4, eT7, eT6, T4).reshape(e78 * t41, e56 * t34)
    U, s, VT = _svd(LT.T @ LB, 1)
    cut = min(p.CHI, e78 * t41, e12 * t12, e56 * t34)
    cut, s_c = _truncate(s, cut, p.Inverse_projector_cut)
    U_c = U[:, :cut] * s_c
    VT_c = VT[:cut] * s_c[:, None]
    PU = (LB @ np.conj(VT_c).T).reshape(e78, t41, cut)
    PL = (LT @ np.conj(U_c)).reshape(e78, t41, cut)
    return PU, PL


def Calc_Next_CTM_single(C1, C4, eT1, eT6, PU, PL):
    """ctm_single.cpp:438-451: C1' = PU (C1 eT1), C4' = (eT6 C4) PL, each
    divided by its largest |entry|."""
    C1n = ein("xtc,xb,bzt->cz", PU, C1, eT1)
    C4n = ein("zbt,bx,xtc->zc", eT6, C4, PL)
    return C1n / max_abs(C1n), C4n / max_abs(C4n)


def Calc_Next_eT_single(eT8, T1, PU, PL):
    """ctm_single.cpp:453-474: the new left edge (cut_PU, cut_PL, t12)."""
    e = ein("yxl,xuc,lurd,ydk->kcr", eT8, PL, T1, PU)
    return e / max_abs(e)


R_RIGHT, R_TOP, R_BOTTOM = (2, 3, 0, 1), (1, 2, 3, 0), (3, 0, 1, 2)


def _projectors(L, args_of, p):
    PUs, PLs = [None] * L, [None] * L
    for n in range(L):
        if not p.CTM_Projector_corner:
            raise NotImplementedError("single-layer up-down projectors are not restated")
        PUs[n], PLs[n] = Calc_projector_left_block_single(*args_of(n), p)
    return PUs, PLs


def Left_move_single(env: Env, A, ix, p, lat):
    """ctm_single.cpp:483-538"""
    L = lat.LY_noskew

    def sites(iy):
        i = lat.index(ix, iy)
        j = lat.right(i)
        k = lat.bottom(j)
        return i, j, k, lat.left(k)

    def args(iy):
        i, j, k, l = sites(iy)
        return env.C1[i], env.C4[l], env.eTt[i], env.eTb[l], env.eTl[l], env.eTl[i], A[i], A[l]

    PUs, PLs = _projectors(L, args, p)
    C1b, C4b, eTlb = list(env.C1), list(env.C4), list(env.eTl)
    for iy in range(L):
        i, j, k, l = sites(iy)
        up, dn = (iy + 1) % L, (iy - 1 + L) % L
        env.C1[j], env.C4[k] = Calc_Next_CTM_single(C1b[i], C4b[l], env.eTt[i], env.eTb[l], PUs[up], PLs[dn])
        env.eTl[j] = Calc_Next_eT_single(eTlb[i], A[i], PUs[iy], PLs[up])
        env.eTl[k] = Calc_Next_eT_single(eTlb[l], A[l], PUs[dn], PLs[iy])


def Right_move_single(env: Env, A, ix, p, lat):
    """ctm_single.cpp:540-603"""
    L = lat.LY_noskew
    Ar = [a.transpose(R_RIGHT) for a in A]

    def sites(iy):
        k = lat.index(ix, iy)
        l = lat.left(k)
        i = lat.top(l)
        return i, lat.right(i), k, l

    def args(iy):
        i, j, k, l = sites(iy)
        return env.C3[k], env.C2[j], env.eTb[k], env.eTt[j], env.eTr[j], env.eTr[k], Ar[k], Ar[j]

    PUs, PLs = _projectors(L, args, p)
    C2b, C3b, eTrb = list(env.C2), list(env.C3), list(env.eTr)
    for iy in range(L):
        i, j, k, l = sites(iy)
        up, dn = (iy + 1) % L, (iy - 1 + L) % L
        env.C3[l], env.C2[i] = Calc_Next_CTM_single(C3b[k], C2b[j], env.eTb[k], env.eTt[j], PUs[dn], PLs[up])
        env.eTr[l] = Calc_Next_eT_single(eTrb[k], Ar[k], PUs[iy], PLs[dn])
        env.eTr[i] = Calc_Next_eT_single(eTrb[j], Ar[j], PUs[up], PLs[iy])


def Top_move_single(env: Env, A, iy, p, lat):
    """ctm_single.cpp:605-668"""
    L = lat.LX_noskew
    At = [a.transpose(R_TOP) for a in A]

    def sites(ix):
        j = lat.index(ix, iy)
        k = lat.bottom(j)
        l = lat.left(k)
        return lat.top(l), j, k, l

    def args(ix):
        i, j, k, l = sites(ix)
        return env.C2[j], env.C1[i], env.eTr[j], env.eTl[i], env.eTt[i], env.eTt[j], At[j], At[i]

    PUs, PLs = _projectors(L, args, p)
    C1b, C2b, eTtb = list(env.C1), list(env.C2), list(env.eTt)
    for ix in range(L):
        i, j, k, l = sites(ix)
        rt, lt = (ix + 1) % L, (ix - 1 + L) % L
        env.C2[k], env.C1[l] = Calc_Next_CTM_single(C2b[j], C1b[i], env.eTr[j], env.eTl[i], PUs[rt], PLs[lt])
        env.eTt[k] = Calc_Next_eT_single(eTtb[j], At[j], PUs[ix], PLs[rt])
        env.eTt[l] = Calc_Next_eT_single(eTtb[i], At[i], PUs[lt], PLs[ix])


def Bottom_move_single(env: Env, A, iy, p, lat):
    """ctm_single.cpp:670-735"""
    L = lat.LX_noskew
    Ab = [a.transpose(R_BOTTOM) for a in A]

    def sites(ix):
        l = lat.index(ix, iy)
        i = lat.top(l)
        j = lat.right(i)
        return i, j, lat.bottom(j), l

    def args(ix):
        i, j, k, l = sites(ix)
        return env.C4[l], env.C3[k], env.eTl[l], env.eTr[k], env.eTb[k], env.eTb[l], Ab[l], Ab[k]

    PUs, PLs = _projectors(L, args, p)
    C3b, C4b, eTbb = list(env.C3), list(env.C4), list(env.eTb)
    for ix in range(L):
        i, j, k, l = sites(ix)
        rt, lt = (ix + 1) % L, (ix - 1 + L) % L
        env.C4[i], env.C3[j] = Calc_Next_CTM_single(C4b[l], C3b[k], env.eTl[l], env.eTr[k], PUs[lt], PLs[rt])
        env.eTb[i] = Calc_Next_eT_single(eTbb[l], Ab[l], PUs[ix], PLs[lt])
        env.eTb[j] = Calc_Next_eT_single(eTbb[k], Ab[k], PUs[rt], PLs[ix])


def initialize_environment_density(A, lat) -> Env:
    """ctm_single.cpp:751-995: corners and edges from the traced tensors
    summed over their outer legs (no padding to chi)."""
    N = lat.N_UNIT
    env = Env(N)
    for i in range(N):
        env.C1[i] = ein("lurd->dr", A[lat.top(lat.left(i))])
        env.C2[i] = ein("lurd->ld", A[lat.top(lat.right(i))])
        env.C3[i] = ein("lurd->ul", A[lat.bottom(lat.right(i))])
        env.C4[i] = ein("lurd->ru", A[lat.bottom(lat.left(i))])
        env.eTt[i] = ein("lurd->lrd", A[lat.top(i)])
        env.eTr[i] = ein("lurd->udl", A[lat.right(i)])
        env.eTb[i] = ein("lurd->rlu", A[lat.bottom(i)])
        env.eTl[i] = ein("lurd->dur", A[lat.left(i)])
    return env


def Calc_CTM_Environment_density(Tn, p, lat, env: Env | None = None):
    """ctm_single.cpp:737-1057: returns (env, count, sig_max)."""
    A = Make_single_tensor_density(Tn)
    if env is None:
        env = initialize_environment_density(A, lat)
    LX, LY = lat.LX_noskew, lat.LY_noskew
    old = {n: list(getattr(env, n)) for n in ("C1", "C2", "C3", "C4")}
    convergence, count, sig_max = False, 0, 0.0
    while (not convergence) and count < p.Max_CTM_Iteration:
        for ix in range(LX):
            Left_move_single(env, A, ix, p, lat)
        for ix in range(0, -LX, -1):
            Right_move_single(env, A, (ix + 1 + LX) % LX, p, lat)
        for iy in range(0, -LY, -1):
            Top_move_single(env, A, (iy + 1 + LY) % LY, p, lat)
        for iy in range(LY):
            Bottom_move_single(env, A, iy, p, lat)
        convergence, sig_max = Check_Convergence_CTM(env, old, p, lat)
        count += 1
        old = {n: list(getattr(env, n)) for n in ("C1", "C2", "C3", "C4")}
    return env, count, sig_max
