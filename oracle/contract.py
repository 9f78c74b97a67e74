"""CTM-environment contractions restated in numpy
(TeNeS src/iTPS/core/contract_itps_ctm/ctm.cpp, ctm_1x1.cpp, ctm_1x2.cpp,
ctm_2x1.cpp, ctm_2x2.cpp, correlation.cpp, src/iTPS/core/full_update.cpp:51-83;
the other cluster shapes of ctm_1x3.cpp ... ctm_4x4.cpp by one row sweep of
double-layer sites, _cluster_rows).  Test infrastructure: the checker of the
GPU library, never part of it."""
from __future__ import annotations

import numpy as np

td = np.tensordot
cj = np.conj


def trace(A, B, axA, axB):
    """mptensor trace(A, B, axA, axB): full contraction to a scalar."""
    return complex(td(A, B, (list(axA), list(axB))))


def Contract_one_site_iTPS_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, Tn1, op1):
    """contract_itps_ctm/ctm.cpp:89-113"""
    return trace(
        td(td(td(td(C2, td(C3, eT2, ([0], [1])), ([1], [1])), td(C1, eT1, ([1], [0])), ([0], [1])),
              td(td(td(C4, eT4, ([1], [0])), eT3, ([0], [1])), cj(Tn1), ([2, 5], [0, 3])),
              ([0, 2, 3, 5], [2, 5, 0, 4])),
           Tn1, ([0, 1, 2, 3], [2, 1, 0, 3])),
        op1, [0, 1], [1, 0])


def Contract_two_sites_horizontal_iTPS_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, Tn1, Tn2,
                                           op1, op2):
    """contract_itps_ctm/ctm.cpp:115-163"""
    rh = td(Tn2, td(td(td(C3, eT4, ([1], [0])), eT3, ([0], [1])), td(cj(Tn2), op2, ([4], [1])),
                    ([2, 5], [3, 2])), ([2, 3, 4], [3, 1, 6]))
    R = td(td(td(C2, eT2, ([0], [1])), rh, ([0, 2, 3], [3, 1, 5])), eT5, ([2], [0]))
    L = td(td(C1, eT1, ([1], [0])), td(C4, eT6, ([1], [0])), ([0], [1]))
    X = td(Tn1, td(L, td(cj(Tn1), R, ([2, 3], [2, 5])), ([0, 2, 3, 5], [3, 1, 5, 0])),
           ([0, 1, 2, 3], [1, 0, 3, 4]))
    return trace(X, op1, [0, 1], [0, 1])


def Contract_two_sites_vertical_iTPS_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, Tn1, Tn2,
                                         op1, op2):
    """contract_itps_ctm/ctm.cpp:165-211"""
    lh = td(Tn2, td(td(td(C4, eT5, ([1], [0])), eT4, ([0], [1])), td(cj(Tn2), op2, ([4], [1])),
                    ([2, 5], [0, 3])), ([0, 3, 4], [1, 3, 6]))
    B = td(td(td(C3, eT3, ([0], [1])), lh, ([0, 2, 3], [3, 1, 5])), eT6, ([2], [0]))
    T = td(td(C2, eT2, ([1], [0])), td(C1, eT1, ([1], [0])), ([0], [1]))
    X = td(Tn1, td(T, td(cj(Tn1), B, ([0, 3], [5, 2])), ([0, 2, 3, 5], [3, 1, 5, 0])),
           ([0, 1, 2, 3], [4, 1, 0, 3]))
    return trace(X, op1, [0, 1], [0, 1])


def Contract_two_sites_horizontal_op12_iTPS_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6,
                                                Tn1, Tn2, op12):
    """contract_itps_ctm/ctm.cpp:213-266"""
    left = td(eT1, td(Tn1, td(cj(Tn1), td(eT5, td(C1, td(C4, eT6, ([1], [0])), ([0], [1])),
                                           ([1], [1])), ([0, 3], [5, 2])), ([0, 3], [6, 4])),
              ([0, 2, 3], [7, 0, 3]))
    right = td(eT2, td(Tn2, td(cj(Tn2), td(eT4, td(C2, td(C3, eT3, ([0], [1])), ([1], [1])),
                                            ([0], [1])), ([2, 3], [5, 2])), ([2, 3], [6, 4])),
               ([1, 2, 3], [7, 1, 4]))
    return trace(op12, td(left, right, ([0, 1, 3, 5], [0, 1, 3, 5])), [0, 1, 2, 3], [0, 2, 1, 3])


def Contract_two_sites_vertical_op12_iTPS_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6,
                                              Tn1, Tn2, op12):
    """contract_itps_ctm/ctm.cpp:268-319"""
    top = td(eT2, td(Tn1, td(cj(Tn1), td(eT6, td(C1, td(C2, eT1, ([0], [1])), ([1], [1])),
                                          ([1], [0])), ([0, 1], [2, 5])), ([0, 1], [4, 6])),
             ([0, 2, 3], [7, 0, 3]))
    bottom = td(eT3, td(Tn2, td(cj(Tn2), td(eT5, td(C3, td(C4, eT4, ([0], [1])), ([1], [1])),
                                             ([0], [1])), ([0, 3], [2, 5])), ([0, 3], [4, 6])),
                ([1, 2, 3], [7, 1, 4]))
    return trace(op12, td(top, bottom, ([0, 1, 3, 5], [0, 1, 3, 5])), [0, 1, 2, 3], [0, 2, 1, 3])


def _c1x1(C, eTt, eTr, eTb, eTl, Tn, op):
    """contract_itps_ctm/ctm_1x1.cpp:27-72"""
    T = Tn[0][0]
    top = td(eTt[0], td(C[0], eTl[0], ([0], [1])), ([0], [0]))
    rest = td(td(C[1], eTr[0], ([1], [0])), td(C[2], td(C[3], eTb[0], ([0], [1])), ([1], [1])),
              ([1], [0]))
    X = td(T, td(top, td(cj(T), rest, ([2, 3], [2, 5])), ([0, 2, 3, 5], [3, 1, 5, 0])),
           ([0, 1, 2, 3], [1, 0, 3, 4]))
    return trace(op[0][0], X, [0, 1], [0, 1])


def _c1x2(C, eTt, eTr, eTb, eTl, Tn, op):
    """contract_itps_ctm/ctm_1x2.cpp:27-87"""
    T0, T1 = Tn[0][0], Tn[0][1]
    renv = td(eTb[1], td(C[1], td(C[2], eTr[0], ([0], [1])), ([1], [1])), ([0], [1]))
    site1 = td(eTt[1], td(T1, td(td(cj(T1), op[0][1], ([4], [1])), renv, ([2, 3], [5, 2])),
                          ([2, 3, 4], [6, 4, 2])), ([1, 2, 3], [5, 1, 3]))
    lower = td(td(C[3], eTb[0], ([0], [1])), site1, ([1], [3]))
    top = td(eTt[0], td(C[0], eTl[0], ([0], [1])), ([0], [0]))
    X = td(T0, td(top, td(cj(T0), lower, ([2, 3], [5, 2])), ([0, 2, 3, 5], [5, 1, 3, 0])),
           ([0, 1, 2, 3], [1, 0, 4, 3]))
    return trace(op[0][0], X, [0, 1], [0, 1])


def _c2x1(C, eTt, eTr, eTb, eTl, Tn, op):
    """contract_itps_ctm/ctm_2x1.cpp:27-87"""
    T0, T1 = Tn[0][0], Tn[1][0]
    benv = td(eTl[1], td(C[2], td(C[3], eTb[0], ([0], [1])), ([1], [1])), ([0], [1]))
    site1 = td(eTr[1], td(T1, td(td(cj(T1), op[1][0], ([4], [1])), benv, ([0, 3], [2, 5])),
                          ([0, 3, 4], [4, 6, 2])), ([1, 2, 3], [5, 1, 3]))
    right = td(td(C[1], eTr[0], ([1], [0])), site1, ([1], [0]))
    top = td(eTt[0], td(C[0], eTl[0], ([0], [1])), ([0], [0]))
    X = td(T0, td(top, td(cj(T0), right, ([2, 3], [2, 4])), ([0, 2, 3, 5], [3, 1, 6, 0])),
           ([0, 1, 2, 3], [1, 0, 3, 4]))
    return trace(op[0][0], X, [0, 1], [0, 1])


def _c2x2(C, eTt, eTr, eTb, eTl, Tn, op):
    """contract_itps_ctm/ctm_2x2.cpp:25-104"""
    T00, T01, T10, T11 = Tn[0][0], Tn[0][1], Tn[1][0], Tn[1][1]
    tr = td(T01, td(td(cj(T01), op[0][1], ([4], [1])), td(eTt[1], td(C[1], eTr[0], ([1], [0])), ([1], [0])),
                    ([1, 2], [2, 5])), ([1, 2, 4], [4, 6, 2]))
    bl = td(T10, td(td(cj(T10), op[1][0], ([4], [1])), td(eTb[0], td(C[3], eTl[1], ([1], [0])), ([1], [0])),
                    ([0, 3], [5, 2])), ([0, 3, 4], [6, 4, 2]))
    br = td(T11, td(td(cj(T11), op[1][1], ([4], [1])), td(eTr[1], td(C[2], eTb[1], ([1], [0])), ([1], [0])),
                    ([2, 3], [2, 5])), ([2, 3, 4], [4, 6, 2]))
    rest = td(tr, td(bl, br, ([1, 3, 4], [0, 2, 5])), ([1, 3, 5], [3, 4, 5]))
    top = td(eTt[0], td(C[0], eTl[0], ([0], [1])), ([0], [0]))
    X = td(T00, td(top, td(cj(T00), rest, ([2, 3], [1, 4])), ([0, 2, 3, 5], [4, 1, 6, 0])),
           ([0, 1, 2, 3], [1, 0, 3, 4]))
    return trace(op[0][0], X, [0, 1], [0, 1])


def _ldot(A, la, B, lb):
    """tensordot of two labelled tensors over every shared label; result
    labels: free labels of A, then of B."""
    sa = [i for i, l in enumerate(la) if l in lb]
    sb = [lb.index(la[i]) for i in sa]
    out = [l for l in la if l not in lb] + [l for l in lb if l not in la]
    return td(A, B, (sa, sb)), out


def _cluster_rows(C, eTt, eTr, eTb, eTl, Tn, op):
    """Any nrow x ncol cluster (the networks of contract_itps_ctm/ctm_NxM.cpp):
    leg orders of ctm.cpp:984-1192, site legs (l, u, r, d, p), the operator's
    first index on the ket, its second on the bra (as in ctm_1x1.cpp:27-72).
    Contracted row by row, top to bottom, with each site fused into its
    double-layer tensor a = T . op . conj(T) first."""
    nr, nc = len(Tn), len(Tn[0])
    B, lb = C[0], [("l", 0), ("t", 0)]
    for c in range(nc):
        B, lb = _ldot(B, lb, eTt[c], [("t", c), ("t", c + 1), ("vk", 0, c), ("vb", 0, c)])
    B, lb = _ldot(B, lb, C[1], [("t", nc), ("r", 0)])
    for r in range(nr):
        B, lb = _ldot(B, lb, eTl[r], [("l", r + 1), ("l", r), ("hk", r, 0), ("hb", r, 0)])
        for c in range(nc):
            a = np.einsum("lurdp,pq,LURDq->lLuUrRdD", Tn[r][c], op[r][c], cj(Tn[r][c]))
            B, lb = _ldot(B, lb, a, [("hk", r, c), ("hb", r, c), ("vk", r, c), ("vb", r, c),
                                     ("hk", r, c + 1), ("hb", r, c + 1),
                                     ("vk", r + 1, c), ("vb", r + 1, c)])
        B, lb = _ldot(B, lb, eTr[r], [("r", r), ("r", r + 1), ("hk", r, nc), ("hb", r, nc)])
    B, lb = _ldot(B, lb, C[3], [("b", 0), ("l", nr)])
    for c in range(nc):
        B, lb = _ldot(B, lb, eTb[c], [("b", c + 1), ("b", c), ("vk", nr, c), ("vb", nr, c)])
    B, lb = _ldot(B, lb, C[2], [("r", nr), ("b", nc)])
    assert not lb
    return complex(B)


def Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, Tn, op):
    """contract_itps_ctm/ctm.cpp:45-87: the 1x1, 1x2, 2x1 and 2x2 trees of
    ctm_1x1/1x2/2x1/2x2.cpp, any other cluster up to 4x4 by _cluster_rows."""
    nrow, ncol = len(Tn), len(Tn[0])
    if (nrow, ncol) == (1, 1):
        return _c1x1(C, eTt, eTr, eTb, eTl, Tn, op)
    if (nrow, ncol) == (1, 2):
        return _c1x2(C, eTt, eTr, eTb, eTl, Tn, op)
    if (nrow, ncol) == (2, 1):
        return _c2x1(C, eTt, eTr, eTb, eTl, Tn, op)
    if (nrow, ncol) == (2, 2):
        return _c2x2(C, eTt, eTr, eTb, eTl, Tn, op)
    if 1 <= nrow <= 4 and 1 <= ncol <= 4:
        return _cluster_rows(C, eTt, eTr, eTb, eTl, Tn, op)
    raise NotImplementedError(f"Contract_{nrow}_{ncol}_iTPS_CTM is not implemented")


def Create_Environment_two_sites(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, Q1, Q2):
    """core/full_update.cpp:51-83; legs (t1, t2, tc1, tc2) of Q1, Q2."""
    rq = td(td(td(td(C3, eT4, ([1], [0])), eT3, ([0], [1])), cj(Q2), ([2, 5], [3, 2])), Q2,
            ([1, 3], [3, 2]))
    right = td(td(C2, eT2, ([0], [1])), rq, ([0, 2, 3], [1, 5, 3]))
    lq = td(td(td(td(C4, eT6, ([1], [0])), eT5, ([0], [1])), cj(Q1), ([2, 5], [0, 2])), Q1,
            ([1, 3], [0, 2]))
    left = td(td(C1, eT1, ([1], [0])), lq, ([0, 2, 3], [0, 4, 2]))
    return td(right, left, ([0, 1], [0, 1])).transpose(3, 1, 2, 0)


def StartCorrelation_iTPS_CTM(C1, C4, eT1, eT3, eT4, Tn1, op):
    """contract_itps_ctm/correlation.cpp:34-58: the left end of a row with the
    operator on Tn1; legs (e1r, e3r, n1r, n2r)."""
    A = td(eT1, td(Tn1, td(td(cj(Tn1), op, ([4], [1])),
                           td(eT3, td(C1, td(C4, eT4, ([1], [0])), ([0], [1])), ([1], [1])),
                           ([0, 3], [5, 2])),
                   ([0, 3, 4], [6, 4, 2])),
           ([0, 2, 3], [5, 0, 2]))
    return A.transpose(0, 3, 1, 2)


def Transfer_iTPS_CTM(A, eT1, eT3, Tn1):
    """contract_itps_ctm/correlation.cpp:60-78: one more column."""
    A = td(eT1, td(Tn1, td(cj(Tn1), td(A, eT3, ([1], [1])), ([0, 3], [2, 5])), ([0, 3, 4], [4, 6, 2])),
           ([0, 2, 3], [4, 0, 2]))
    return A.transpose(0, 3, 1, 2)


def FinishCorrelation_iTPS_CTM(A, C2, C3, eT1, eT2, eT3, Tn1, op):
    """contract_itps_ctm/correlation.cpp:80-105: the right end with the
    operator on Tn1."""
    return trace(op, td(Tn1, td(td(A, eT1, ([0], [0])),
                                td(cj(Tn1), td(eT3, td(C2, td(C3, eT2, ([0], [1])), ([1], [1])), ([0], [1])),
                                   ([2, 3], [5, 2])),
                                ([0, 2, 3, 5], [3, 0, 5, 1])),
                        ([0, 1, 2, 3], [0, 1, 4, 3])),
                 [0, 1], [0, 1])
