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


def _truncate_e78(s, cut0, e78, eps):
    """ctm_single.cpp:396-415, the full-SVD branch of the up-down projectors:
    s_c starts as cut0 zeros and the loop runs over the first e78 values
    (not the first cut0); the first value at or below eps (relative to s[0])
    becomes the cut, which can exceed cut0.  Entries from cut0 on are zero
    (the vector grows by value-initialised entries) and so are entries from
    e78 on.  Returns (cut, s_c)."""
    brk = None
    for i in range(min(e78, s.size)):
        if not (s[i] / s[0] > eps):
            brk = i
            break
    cut = cut0 if brk is None else brk
    s_c = np.zeros(cut)
    m = min(cut, cut0, e78)
    s_c[:m] = 1.0 / np.sqrt(s[:m])
    return cut, s_c


def Calc_projector_updown_blocks_single(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, eT7, eT8,
                                        T1, T2, T3, T4, p: PEPSParameters):
    """ctm_single.cpp:267-436 with the exact SVD in both branches (the RSVD
    branch truncates as the double layer, the full-SVD branch with the e78
    loop of _truncate_e78).  Returns PU, PL with legs (e78, t41, cut)."""
    e12, e34, e56, e78 = eT1.shape[1], eT3.shape[1], eT5.shape[1], eT7.shape[1]
    t12, t23, t34, t41 = T1.shape[2], T2.shape[3], T3.shape[0], T4.shape[1]
    if t41 == 1:
        I = np.eye(e78, dtype=C1.dtype)
        return I.reshape(e78, 1, e78), I.reshape(e78, 1, e78)
    # the four enlarged corners (ctm_single.cpp:287-344)
    LT = ein("ab,bxu,yal,lurd->ydxr", C1, eT1, eT8, T1).reshape(e78 * t41, e12 * t12)
    RT = ein("ab,bxr,yau,lurd->ylxd", C2, eT3, eT2, T2).reshape(e12 * t12, e34 * t23)
    RB = ein("ab,bxd,yar,lurd->xlyu", C3, eT5, eT4, T3).reshape(e56 * t34, e34 * t23)
    LB = ein("ab,bxl,yad,lurd->xuyr", C4, eT7, eT6, T4).reshape(e78 * t41, e56 * t34)
    R1 = RT.T @ LT.T  # (e34 t23, e78 t41)
    R2 = RB.T @ LB.T
    U, s, VT = np.linalg.svd(R1 @ R2.T)
    cut0 = min(p.CHI, e34 * t23)
    if t23 != 1 and p.Use_RSVD:
        cut, s_c = _truncate(s, cut0, p.Inverse_projector_cut)
    else:
        cut, s_c = _truncate_e78(s, cut0, e78, p.Inverse_projector_cut)
    assert cut <= U.shape[1]
    U_c = U[:, :cut] * s_c
    VT_c = VT[:cut] * s_c[:, None]
    PU = (R2.T @ np.conj(VT_c).T).reshape(e78, t41, cut)
    PL = (R1.T @ np.conj(U_c)).reshape(e78, t41, cut)
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


def _projectors(L, args_of, ud_args_of, p):
    PUs, PLs = [None] * L, [None] * L
    for n in range(L):
        if p.CTM_Projector_corner:
            PUs[n], PLs[n] = Calc_projector_left_block_single(*args_of(n), p)
        else:
            PUs[n], PLs[n] = Calc_projector_updown_blocks_single(*ud_args_of(n), p)
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

    def ud_args(iy):
        i, j, k, l = sites(iy)
        return (env.C1[i], env.C2[j], env.C3[k], env.C4[l], env.eTt[i], env.eTt[j], env.eTr[j],
                env.eTr[k], env.eTb[k], env.eTb[l], env.eTl[l], env.eTl[i], A[i], A[j], A[k], A[l])

    PUs, PLs = _projectors(L, args, ud_args, p)
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

    def ud_args(iy):
        i, j, k, l = sites(iy)
        return (env.C3[k], env.C4[l], env.C1[i], env.C2[j], env.eTb[k], env.eTb[l], env.eTl[l],
                env.eTl[i], env.eTt[i], env.eTt[j], env.eTr[j], env.eTr[k], Ar[k], Ar[l], Ar[i], Ar[j])

    PUs, PLs = _projectors(L, args, ud_args, p)
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

    def ud_args(ix):
        i, j, k, l = sites(ix)
        return (env.C2[j], env.C3[k], env.C4[l], env.C1[i], env.eTr[j], env.eTr[k], env.eTb[k],
                env.eTb[l], env.eTl[l], env.eTl[i], env.eTt[i], env.eTt[j], At[j], At[k], At[l], At[i])

    PUs, PLs = _projectors(L, args, ud_args, p)
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

    def ud_args(ix):
        i, j, k, l = sites(ix)
        return (env.C4[l], env.C1[i], env.C2[j], env.C3[k], env.eTl[l], env.eTl[i], env.eTt[i],
                env.eTt[j], env.eTr[j], env.eTr[k], env.eTb[k], env.eTb[l], Ab[l], Ab[i], Ab[j], Ab[k])

    PUs, PLs = _projectors(L, args, ud_args, p)
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
