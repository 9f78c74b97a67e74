"""Single-layer environment contractions of an iTPO restated in numpy (TeNeS
src/iTPS/core/contract_density_ctm/ctm.cpp, ctm_NxM.cpp, correlation.cpp).
Test infrastructure: the checker of the GPU library, never part of it.

TPO site tensors (left, up, right, down, phys_ket, phys_bra); a one-site
operator op[a, b] meets phys_ket with a and phys_bra with b; a two-site
operator op12[a1, a2, b1, b2] meets site 1 with (a1, b1) and site 2 with
(a2, b2).  Edges (chi, chi, D) in the iTPS leg order minus the bra leg.
The functions also carry the iTPS names of oracle.contract, so the
measurement drivers of oracle.itps run on either network.
"""
from __future__ import annotations

import numpy as np

from .contract import _ldot

ein = np.einsum


def Contract_one_site_density_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, Tn1, op1):
    """contract_density_ctm/ctm.cpp:87-112: Tr(rho op) on one site with its
    environment (eT1..eT4 = top, right, bottom, left edges)."""
    top = ein("ab,bcu,cd,dex,hay->ehuxy", C1, eT1, C2, eT2, eT4, optimize="optimal")
    bot = ein("ef,fgv,gh->ehv", C3, eT3, C4, optimize="optimal")
    return complex(ein("ehuxy,ehv,yuxvpq,pq->", top, bot, Tn1, op1, optimize="optimal"))


def Contract_two_sites_horizontal_op12_density_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6,
                                                   Tn1, Tn2, op12):
    """contract_density_ctm/ctm.cpp:195-233: Tn1 left, Tn2 right; eT1, eT2
    the top edges, eT3 the right, eT4, eT5 the bottom edges (right site
    first), eT6 the left.  Left and right halves first, then op12."""
    L = ein("ab,bcu,hiv,ij,jay,yumvpq->chmpq", C1, eT1, eT5, C4, eT6, Tn1, optimize="optimal")
    R = ein("cdw,de,efx,fg,ghz,mwxzrt->chmrt", eT2, C2, eT3, C3, eT4, Tn2, optimize="optimal")
    return complex(ein("chmpq,chmrt,prqt->", L, R, op12, optimize="optimal"))


def Contract_two_sites_vertical_op12_density_CTM(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6,
                                                 Tn1, Tn2, op12):
    """contract_density_ctm/ctm.cpp:235-273: Tn1 top, Tn2 bottom; eT1 the top
    edge, eT2, eT3 the right edges, eT4 the bottom, eT5, eT6 the left edges
    (bottom site first).  Upper and lower halves first, then op12."""
    U = ein("ab,bcu,cd,dex,jay,yuxmpq->ejmpq", C1, eT1, C2, eT2, eT6, Tn1, optimize="optimal")
    D = ein("efw,fg,ghv,hi,ijz,zmwvrt->ejmrt", eT3, C3, eT4, C4, eT5, Tn2, optimize="optimal")
    return complex(ein("ejmpq,ejmrt,prqt->", U, D, op12, optimize="optimal"))


def Contract_density_CTM(C, eTt, eTr, eTb, eTl, Tn, op):
    """contract_density_ctm/ctm.cpp:45-85 (the ctm_NxM.cpp networks, up to
    4x4): row by row, each site first traced with its operator."""
    nr, nc = len(Tn), len(Tn[0])
    if not (1 <= nr <= 4 and 1 <= nc <= 4):
        raise NotImplementedError(f"Contract_density_{nr}_{nc} is not implemented")
    B, lb = C[0], [("l", 0), ("t", 0)]
    for c in range(nc):
        B, lb = _ldot(B, lb, eTt[c], [("t", c), ("t", c + 1), ("v", 0, c)])
    B, lb = _ldot(B, lb, C[1], [("t", nc), ("r", 0)])
    for r in range(nr):
        B, lb = _ldot(B, lb, eTl[r], [("l", r + 1), ("l", r), ("h", r, 0)])
        for c in range(nc):
            a = ein("lurdpq,pq->lurd", Tn[r][c], op[r][c])
            B, lb = _ldot(B, lb, a, [("h", r, c), ("v", r, c), ("h", r, c + 1), ("v", r + 1, c)])
        B, lb = _ldot(B, lb, eTr[r], [("r", r), ("r", r + 1), ("h", r, nc)])
    B, lb = _ldot(B, lb, C[3], [("b", 0), ("l", nr)])
    for c in range(nc):
        B, lb = _ldot(B, lb, eTb[c], [("b", c + 1), ("b", c), ("v", nr, c)])
    B, lb = _ldot(B, lb, C[2], [("r", nr), ("b", nc)])
    assert not lb
    return complex(B)


def StartCorrelation_density_CTM(C1, C4, eT1, eT3, eT4, Tn1, op):
    """contract_density_ctm/correlation.cpp:34-52: the left end of a row with
    the operator on Tn1; legs (eT1 right, eT3 right, Tn1 right)."""
    a = ein("lurdpq,pq->lurd", Tn1, op)
    col = ein("ax,xtu,yal->tuyl", C1, eT1, eT4, optimize="optimal")
    col = ein("tuyl,zy,bzd->tulbd", col, C4, eT3, optimize="optimal")
    return ein("tulbd,lurd->tbr", col, a, optimize="optimal")


def Transfer_density_CTM(A, eT1, eT3, Tn1):
    """contract_density_ctm/correlation.cpp:54-70: one more column."""
    a = ein("lurdpp->lurd", Tn1)
    return ein("tbr,tsu,cbd,rued->sce", A, eT1, eT3, a, optimize="optimal")


def FinishCorrelation_density_CTM(A, C2, C3, eT1, eT2, eT3, Tn1, op):
    """contract_density_ctm/correlation.cpp:72-95: closes the row with the
    operator on Tn1."""
    a = ein("ruxdpq,pq->ruxd", Tn1, op)
    right = ein("se,efx,fg,gbd->sxbd", C2, eT2, C3, eT3, optimize="optimal")
    left = ein("tbr,tsu->bsru", A, eT1, optimize="optimal")
    return complex(ein("bsru,sxbd,ruxd->", left, right, a, optimize="optimal"))


# iTPS names, for the shared measurement drivers
Contract_one_site_iTPS_CTM = Contract_one_site_density_CTM
Contract_two_sites_horizontal_op12_iTPS_CTM = Contract_two_sites_horizontal_op12_density_CTM
Contract_two_sites_vertical_op12_iTPS_CTM = Contract_two_sites_vertical_op12_density_CTM
Contract_iTPS_CTM = Contract_density_CTM
StartCorrelation_iTPS_CTM = StartCorrelation_density_CTM
Transfer_iTPS_CTM = Transfer_density_CTM
FinishCorrelation_iTPS_CTM = FinishCorrelation_density_CTM
