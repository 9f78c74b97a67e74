"""CTMRG restated in numpy (TeNeS src/iTPS/core/ctm.cpp), real or complex.

Follows the reference's contraction order, dynamic bond dimension (the
projector cut can shrink chi) and control flow; mptensor's tensordot /
transpose / svd / max_abs are numpy tensordot / transpose / linalg.svd /
max |x|.  Tensor leg conventions:
  Tn (left, up, right, down, phys); C (chi, chi); eT (chi, chi, D, D).
"""
from __future__ import annotations

import numpy as np

td = np.tensordot


class PEPSParameters:
    """CTM fields of PEPS_Parameters (PEPS_Parameters.cpp:29-77 defaults)."""

    def __init__(self, **kw):
        self.CHI = 2
        self.Inverse_projector_cut = 1e-12
        self.CTM_Convergence_Epsilon = 1e-6
        self.Max_CTM_Iteration = 100
        self.CTM_Projector_corner = True
        self.Use_RSVD = False
        self.RSVD_Oversampling_factor = 2.0
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)


def max_abs(T: np.ndarray) -> float:
    return float(np.max(np.abs(T)))


def _svd(M: np.ndarray, nrow_legs: int):
    shp = M.shape
    r = int(np.prod(shp[:nrow_legs]))
    U, s, Vh = np.linalg.svd(M.reshape(r, -1), full_matrices=False)
    k = s.size
    return U.reshape(shp[:nrow_legs] + (k,)), s, Vh.reshape((k,) + shp[nrow_legs:])


def _truncate(s: np.ndarray, cut: int, eps: float):
    """ctm.cpp:275-288: keep s[i]/s[0] > eps, returns (cut, 1/sqrt(s))."""
    denom = s[0]
    s_c = []
    for i in range(cut):
        if s[i] / denom > eps:
            s_c.append(1.0 / np.sqrt(s[i]))
        else:
            cut = i
            break
    return cut, np.array(s_c)


def _identity_projector(e78, t41, dtype):
    I = np.eye(e78, dtype=dtype)
    return I.reshape(e78, t41, t41, e78), I.reshape(e78, t41, t41, e78)


def enlarged_LT(C1, eT1, eT8, Tn1):
    """ctm.cpp:207-212, legs (t12, t41, e12, e78, tc12, tc41)."""
    return td(Tn1, td(td(td(C1, eT1, ([1], [0])), eT8, ([0], [1])), np.conj(Tn1), ([2, 5], [1, 0])),
              ([0, 1, 4], [3, 1, 6]))


def enlarged_LB(C4, eT7, eT6, Tn4):
    """ctm.cpp:222-227, legs (t41, t34, e78, e56, tc41, tc34)."""
    return td(Tn4, td(td(td(C4, eT7, ([1], [0])), eT6, ([0], [1])), np.conj(Tn4), ([2, 5], [0, 3])),
              ([0, 3, 4], [1, 3, 6]))


def Calc_projector_left_block(C1, C4, eT1, eT6, eT7, eT8, Tn1, Tn4, p: PEPSParameters):
    """ctm.cpp:174-318 (full-SVD branch; the RSVD branch uses the exact SVD
    here as well, see the package docstring)."""
    e12 = eT1.shape[1]
    e56 = eT6.shape[0]
    e78 = eT8.shape[0]
    t12 = Tn1.shape[2]
    t41 = Tn1.shape[3]
    t34 = Tn4.shape[2]
    if t41 == 1:
        return _identity_projector(e78, t41, C1.dtype)
    LT = enlarged_LT(C1, eT1, eT8, Tn1)
    LB = enlarged_LB(C4, eT7, eT6, Tn4)
    rsvd = (t12 != 1 and t34 != 1) and p.Use_RSVD
    U, s, VT = _svd(td(LT, LB, ([1, 3, 5], [0, 2, 4])), 3)
    if rsvd:
        cut = min(p.CHI, e78 * t41 * t41, e12 * t12 * t12, e56 * t34 * t34)
    else:
        cut = min(s.size, p.CHI)
    cut, s_c = _truncate(s, cut, p.Inverse_projector_cut)
    U_c = U[..., :cut] * s_c
    VT_c = VT[:cut] * s_c[:, None, None, None]
    PU = td(LB, np.conj(VT_c), ([1, 3, 5], [1, 2, 3])).transpose(1, 0, 2, 3)
    PL = td(LT, np.conj(U_c), ([0, 2, 4], [0, 1, 2])).transpose(1, 0, 2, 3)
    return PU, PL


def Calc_projector_updown_blocks(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, eT7, eT8,
                                 Tn1, Tn2, Tn3, Tn4, p: PEPSParameters):
    """ctm.cpp:320-491 (full-SVD branch)."""
    e34 = eT3.shape[1]
    e78 = eT7.shape[1]
    t23 = Tn2.shape[3]
    t41 = Tn4.shape[1]
    if t41 == 1:
        return _identity_projector(e78, t41, C1.dtype)
    LT = enlarged_LT(C1, eT1, eT8, Tn1)
    RT = td(Tn2, td(td(td(C2, eT3, ([1], [0])), eT2, ([0], [1])), np.conj(Tn2), ([2, 5], [2, 1])),
            ([1, 2, 4], [3, 1, 6]))
    RB = td(Tn3, td(td(td(C3, eT5, ([1], [0])), eT4, ([0], [1])), np.conj(Tn3), ([2, 5], [3, 2])),
            ([2, 3, 4], [3, 1, 6]))
    LB = enlarged_LB(C4, eT7, eT6, Tn4)
    R1 = td(RT, LT, ([0, 3, 4], [0, 2, 4]))
    R2 = td(RB, LB, ([0, 2, 4], [1, 3, 5]))
    U, s, VT = _svd(td(R1, R2, ([3, 4, 5], [3, 4, 5])), 3)
    rsvd = t23 != 1 and p.Use_RSVD
    cut = min(p.CHI, e34 * t23 * t23) if rsvd else min(s.size, p.CHI)
    cut, s_c = _truncate(s, cut, p.Inverse_projector_cut)
    U_c = U[..., :cut] * s_c
    VT_c = VT[:cut] * s_c[:, None, None, None]
    PU = td(R2, np.conj(VT_c), ([0, 1, 2], [1, 2, 3])).transpose(1, 0, 2, 3)
    PL = td(R1, np.conj(U_c), ([0, 1, 2], [0, 1, 2])).transpose(1, 0, 2, 3)
    return PU, PL


def Calc_Next_CTM(C1, C4, eT1, eT6, PU, PL):
    """ctm.cpp:493-518"""
    C1_out = td(PU, td(C1, eT1, ([1], [0])), ([0, 1, 2], [0, 2, 3]))
    C4_out = td(td(eT6, C4, ([1], [0])), PL, ([3, 1, 2], [0, 1, 2]))
    return C1_out / max_abs(C1_out), C4_out / max_abs(C4_out)


def Calc_Next_eT(eT8, Tn1, PU, PL):
    """ctm.cpp:520-552"""
    e = td(td(Tn1, td(np.conj(Tn1), td(eT8, PL, ([1], [0])), ([0, 1], [2, 4])), ([0, 1, 4], [4, 5, 2])),
           PU, ([1, 3, 4], [1, 2, 0])).transpose(3, 2, 0, 1)
    return e / max_abs(e)


class Env:
    """The environment arrays of iTPS (iTPS.hpp:293-315)."""

    def __init__(self, N):
        self.C1 = [None] * N
        self.C2 = [None] * N
        self.C3 = [None] * N
        self.C4 = [None] * N
        self.eTt = [None] * N
        self.eTr = [None] * N
        self.eTb = [None] * N
        self.eTl = [None] * N


def _projector(env, Tn, p, corner_args, updown_args, rot):
    t = [np.transpose(Tn[i], rot) if rot else Tn[i] for i in range(len(Tn))]
    if p.CTM_Projector_corner:
        Cs, eTs, T1, T4 = corner_args
        return Calc_projector_left_block(*Cs, *eTs, t[T1], t[T4], p)
    Cs, eTs, Ts = updown_args
    return Calc_projector_updown_blocks(*Cs, *eTs, *[t[i] for i in Ts], p)


def Left_move(env: Env, Tn, ix, p, lat):
    """ctm.cpp:562-616"""
    L = lat.LY_noskew
    PUs, PLs = [None] * L, [None] * L
    for iy in range(L):
        i = lat.index(ix, iy)
        j = lat.right(i)
        k = lat.bottom(j)
        l = lat.left(k)
        if p.CTM_Projector_corner:
            PUs[iy], PLs[iy] = Calc_projector_left_block(
                env.C1[i], env.C4[l], env.eTt[i], env.eTb[l], env.eTl[l], env.eTl[i], Tn[i], Tn[l], p)
        else:
            PUs[iy], PLs[iy] = Calc_projector_updown_blocks(
                env.C1[i], env.C2[j], env.C3[k], env.C4[l], env.eTt[i], env.eTt[j], env.eTr[j],
                env.eTr[k], env.eTb[k], env.eTb[l], env.eTl[l], env.eTl[i], Tn[i], Tn[j], Tn[k], Tn[l], p)
    C1b, C4b, eTlb = list(env.C1), list(env.C4), list(env.eTl)
    for iy in range(L):
        i = lat.index(ix, iy)
        j = lat.right(i)
        k = lat.bottom(j)
        l = lat.left(k)
        up = (iy + 1) % L
        dn = (iy - 1 + L) % L
        env.C1[j], env.C4[k] = Calc_Next_CTM(C1b[i], C4b[l], env.eTt[i], env.eTb[l], PUs[up], PLs[dn])
        env.eTl[j] = Calc_Next_eT(eTlb[i], Tn[i], PUs[iy], PLs[up])
        env.eTl[k] = Calc_Next_eT(eTlb[l], Tn[l], PUs[dn], PLs[iy])


def Right_move(env: Env, Tn, ix, p, lat):
    """ctm.cpp:618-681"""
    L = lat.LY_noskew
    r = (2, 3, 0, 1, 4)
    PUs, PLs = [None] * L, [None] * L
    for iy in range(L):
        k = lat.index(ix, iy)
        l = lat.left(k)
        i = lat.top(l)
        j = lat.right(i)
        if p.CTM_Projector_corner:
            PUs[iy], PLs[iy] = Calc_projector_left_block(
                env.C3[k], env.C2[j], env.eTb[k], env.eTt[j], env.eTr[j], env.eTr[k],
                Tn[k].transpose(r), Tn[j].transpose(r), p)
        else:
            PUs[iy], PLs[iy] = Calc_projector_updown_blocks(
                env.C3[k], env.C4[l], env.C1[i], env.C2[j], env.eTb[k], env.eTb[l], env.eTl[l],
                env.eTl[i], env.eTt[i], env.eTt[j], env.eTr[j], env.eTr[k], Tn[k].transpose(r),
                Tn[l].transpose(r), Tn[i].transpose(r), Tn[j].transpose(r), p)
    C2b, C3b, eTrb = list(env.C2), list(env.C3), list(env.eTr)
    for iy in range(L):
        k = lat.index(ix, iy)
        l = lat.left(k)
        i = lat.top(l)
        j = lat.right(i)
        up = (iy + 1) % L
        dn = (iy - 1 + L) % L
        env.C3[l], env.C2[i] = Calc_Next_CTM(C3b[k], C2b[j], env.eTb[k], env.eTt[j], PUs[dn], PLs[up])
        env.eTr[l] = Calc_Next_eT(eTrb[k], Tn[k].transpose(r), PUs[iy], PLs[dn])
        env.eTr[i] = Calc_Next_eT(eTrb[j], Tn[j].transpose(r), PUs[up], PLs[iy])


def Top_move(env: Env, Tn, iy, p, lat):
    """ctm.cpp:683-746"""
    L = lat.LX_noskew
    r = (1, 2, 3, 0, 4)
    PUs, PLs = [None] * L, [None] * L
    for ix in range(L):
        j = lat.index(ix, iy)
        k = lat.bottom(j)
        l = lat.left(k)
        i = lat.top(l)
        if p.CTM_Projector_corner:
            PUs[ix], PLs[ix] = Calc_projector_left_block(
                env.C2[j], env.C1[i], env.eTr[j], env.eTl[i], env.eTt[i], env.eTt[j],
                Tn[j].transpose(r), Tn[i].transpose(r), p)
        else:
            PUs[ix], PLs[ix] = Calc_projector_updown_blocks(
                env.C2[j], env.C3[k], env.C4[l], env.C1[i], env.eTr[j], env.eTr[k], env.eTb[k],
                env.eTb[l], env.eTl[l], env.eTl[i], env.eTt[i], env.eTt[j], Tn[j].transpose(r),
                Tn[k].transpose(r), Tn[l].transpose(r), Tn[i].transpose(r), p)
    C1b, C2b, eTtb = list(env.C1), list(env.C2), list(env.eTt)
    for ix in range(L):
        j = lat.index(ix, iy)
        k = lat.bottom(j)
        l = lat.left(k)
        i = lat.top(l)
        rt = (ix + 1) % L
        lf = (ix - 1 + L) % L
        env.C2[k], env.C1[l] = Calc_Next_CTM(C2b[j], C1b[i], env.eTr[j], env.eTl[i], PUs[rt], PLs[lf])
        env.eTt[k] = Calc_Next_eT(eTtb[j], Tn[j].transpose(r), PUs[ix], PLs[rt])
        env.eTt[l] = Calc_Next_eT(eTtb[i], Tn[i].transpose(r), PUs[lf], PLs[ix])


def Bottom_move(env: Env, Tn, iy, p, lat):
    """ctm.cpp:748-813"""
    L = lat.LX_noskew
    r = (3, 0, 1, 2, 4)
    PUs, PLs = [None] * L, [None] * L
    for ix in range(L):
        l = lat.index(ix, iy)
        i = lat.top(l)
        j = lat.right(i)
        k = lat.bottom(j)
        if p.CTM_Projector_corner:
            PUs[ix], PLs[ix] = Calc_projector_left_block(
                env.C4[l], env.C3[k], env.eTl[l], env.eTr[k], env.eTb[k], env.eTb[l],
                Tn[l].transpose(r), Tn[k].transpose(r), p)
        else:
            PUs[ix], PLs[ix] = Calc_projector_updown_blocks(
                env.C4[l], env.C1[i], env.C2[j], env.C3[k], env.eTl[l], env.eTl[i], env.eTt[i],
                env.eTt[j], env.eTr[j], env.eTr[k], env.eTb[k], env.eTb[l], Tn[l].transpose(r),
                Tn[i].transpose(r), Tn[j].transpose(r), Tn[k].transpose(r), p)
    C3b, C4b, eTbb = list(env.C3), list(env.C4), list(env.eTb)
    for ix in range(L):
        l = lat.index(ix, iy)
        i = lat.top(l)
        j = lat.right(i)
        k = lat.bottom(j)
        rt = (ix + 1) % L
        lf = (ix - 1 + L) % L
        env.C4[i], env.C3[j] = Calc_Next_CTM(C4b[l], C3b[k], env.eTl[l], env.eTr[k], PUs[lf], PLs[rt])
        env.eTb[i] = Calc_Next_eT(eTbb[l], Tn[l].transpose(r), PUs[ix], PLs[lf])
        env.eTb[j] = Calc_Next_eT(eTbb[k], Tn[k].transpose(r), PUs[rt], PLs[ix])


def normalized_spectrum(C) -> np.ndarray:
    s = np.linalg.svd(C, compute_uv=False)
    return s / np.sqrt(np.sum(s * s))


def Check_Convergence_CTM(env: Env, old, p, lat):
    """ctm.cpp:815-969 (same early exit; a missing old entry counts as 0)."""
    sig_max = 0.0
    for i in range(lat.N_UNIT):
        for name in ("C1", "C2", "C3", "C4"):
            ln = normalized_spectrum(getattr(env, name)[i])
            lo = normalized_spectrum(old[name][i])
            lo2 = np.zeros_like(ln)
            m = min(ln.size, lo.size)
            lo2[:m] = lo[:m]
            sig = float(np.sqrt(np.sum((ln - lo2) ** 2)))
            if sig > p.CTM_Convergence_Epsilon:
                return False, sig
            if sig > sig_max:
                sig_max = sig
    return True, sig_max


def _fit(T, chi):
    """extend/slice the two chi legs to chi (ctm.cpp:998-1012)."""
    d1, d2 = T.shape[0], T.shape[1]
    T = T[: min(d1, chi), : min(d2, chi)]
    if T.shape[0] < chi or T.shape[1] < chi:
        pad = [(0, chi - T.shape[0]), (0, chi - T.shape[1])] + [(0, 0)] * (T.ndim - 2)
        T = np.pad(T, pad)
    return T


def initialize_environment(Tn, p, lat) -> Env:
    """ctm.cpp:984-1193.  Merged legs (a, a*) are ordered row-major."""
    env = Env(lat.N_UNIT)
    chi = p.CHI
    c = np.conj
    for i in range(lat.N_UNIT):
        T = Tn[lat.top(lat.left(i))]
        env.C1[i] = _fit(td(T, c(T), ([0, 1, 4], [0, 1, 4])).transpose(1, 3, 0, 2)
                         .reshape(T.shape[3] ** 2, T.shape[2] ** 2), chi)
        T = Tn[lat.top(lat.right(i))]
        env.C2[i] = _fit(td(T, c(T), ([1, 2, 4], [1, 2, 4])).transpose(0, 2, 1, 3)
                         .reshape(T.shape[0] ** 2, T.shape[3] ** 2), chi)
        T = Tn[lat.bottom(lat.right(i))]
        env.C3[i] = _fit(td(T, c(T), ([2, 3, 4], [2, 3, 4])).transpose(1, 3, 0, 2)
                         .reshape(T.shape[1] ** 2, T.shape[0] ** 2), chi)
        T = Tn[lat.bottom(lat.left(i))]
        env.C4[i] = _fit(td(T, c(T), ([0, 3, 4], [0, 3, 4])).transpose(1, 3, 0, 2)
                         .reshape(T.shape[2] ** 2, T.shape[1] ** 2), chi)
        T = Tn[lat.top(i)]
        env.eTt[i] = _fit(td(T, c(T), ([1, 4], [1, 4])).transpose(0, 3, 1, 4, 2, 5)
                          .reshape(T.shape[0] ** 2, T.shape[2] ** 2, T.shape[3], T.shape[3]), chi)
        T = Tn[lat.right(i)]
        env.eTr[i] = _fit(td(T, c(T), ([2, 4], [2, 4])).transpose(1, 4, 2, 5, 0, 3)
                          .reshape(T.shape[1] ** 2, T.shape[3] ** 2, T.shape[0], T.shape[0]), chi)
        T = Tn[lat.bottom(i)]
        env.eTb[i] = _fit(td(T, c(T), ([3, 4], [3, 4])).transpose(2, 5, 0, 3, 1, 4)
                          .reshape(T.shape[2] ** 2, T.shape[0] ** 2, T.shape[1], T.shape[1]), chi)
        T = Tn[lat.left(i)]
        env.eTl[i] = _fit(td(T, c(T), ([0, 4], [0, 4])).transpose(2, 5, 0, 3, 1, 4)
                          .reshape(T.shape[3] ** 2, T.shape[1] ** 2, T.shape[2], T.shape[2]), chi)
    return env


def sweep(env: Env, Tn, p, lat):
    """One iteration of the loop body of ctm.cpp:1203-1228."""
    LX, LY = lat.LX_noskew, lat.LY_noskew
    for ix in range(LX):
        Left_move(env, Tn, ix, p, lat)
    for ix in range(0, -LX, -1):
        Right_move(env, Tn, (ix + 1 + LX) % LX, p, lat)
    for iy in range(0, -LY, -1):
        Top_move(env, Tn, (iy + 1 + LY) % LY, p, lat)
    for iy in range(LY):
        Bottom_move(env, Tn, iy, p, lat)


def Calc_CTM_Environment(Tn, p, lat, env: Env | None = None, verbose=False):
    """ctm.cpp:971-1253; returns (env, count, sig_max)."""
    if env is None:
        env = initialize_environment(Tn, p, lat)
    convergence = False
    count = 0
    old = {n: list(getattr(env, n)) for n in ("C1", "C2", "C3", "C4")}
    sig_max = 0.0
    while (not convergence) and count < p.Max_CTM_Iteration:
        sweep(env, Tn, p, lat)
        convergence, sig_max = Check_Convergence_CTM(env, old, p, lat)
        count += 1
        old = {n: list(getattr(env, n)) for n in ("C1", "C2", "C3", "C4")}
        if verbose:
            print("CTM: count, sig_max", count, sig_max)
    return env, count, sig_max
