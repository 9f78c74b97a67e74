"""Oracle (test infrastructure, not product): the bond energy of an iTPS with
its CTM environment held fixed, and the gradient of that energy with respect
to every site tensor.

The value of one nearest-neighbour bond is the op12 network of
Contract_two_sites_{horizontal,vertical}_op12_iTPS_CTM
(src/iTPS/core/contract_itps_ctm/ctm.cpp:213-319) divided by the cluster norm
(twosite_obs.cpp:168-262, density.cpp:39-76 for the sum over bonds).  Here the
network of one bond is written with explicit einsum labels from the TeNeS leg
conventions (C: (in, out) around the clockwise ring; eT: (in, out, D, D*);
Tn: (left, up, right, down, phys)) and split into two halves, one per site,
whose product over the shared legs is the two-site reduced density matrix
rho(ket1, bra1, ket2, bra2); the value is tr(rho op).  Tests pin the value
against oracle/contract.py (which the reference's goldens pin) and the
gradient against central finite differences of the energy.

A real site tensor enters a bond twice (ket and bra); its gradient is the sum
of the two networks with that occurrence removed.  The gradient has no
counterpart in the reference (TeNeS optimises by imaginary-time evolution);
it is the adjoint asked of this path.
"""
from __future__ import annotations

import numpy as np

from .itps import cluster

_es = np.einsum

# Each half: environment steps (einsum specs applied left to right to the
# running tensor), then the ket and bra site tensors.  Labels as in the full
# horizontal network
#   C1 ab, eT1 bckK, eT2 cdrR, C2 de, eT3 efmM, C3 fg, eT4 ghnN, eT5 hisS,
#   C4 ij, eT6 jalL, Tn1 lkxsp, cTn1 LKXSq, Tn2 xrmnt, cTn2 XRMNu
# and the vertical one
#   C1 ab, eT1 bckK, C2 cd, eT2 demM, eT3 efrR, C3 fg, eT4 ghnN, C4 hi,
#   eT5 ijsS, eT6 jalL, Tn1 lkmyp, cTn1 LKMYq, Tn2 syrnt, cTn2 SYRNu.
_HALVES = {
    "h1": (["ab,jalL->bjlL", "bjlL,ij->bilL", "bilL,hisS->bhlLsS", "bhlLsS,bckK->chlLsSkK"],
           "lkxsp", "LKXSq", "chlLsSkK", "chLSKxp", "chxpXq"),
    "h2": (["de,efmM->dfmM", "dfmM,fg->dgmM", "dgmM,ghnN->dhmMnN", "dhmMnN,cdrR->chmMnNrR"],
           "xrmnt", "XRMNu", "chmMnNrR", "chMNRxt", "chxtXu"),
    "v1": (["ab,bckK->ackK", "ackK,cd->adkK", "adkK,demM->aekKmM", "aekKmM,jalL->ejkKmMlL"],
           "lkmyp", "LKMYq", "ejkKmMlL", "ejKMLyp", "ejypYq"),
    "v2": (["fg,ghnN->fhnN", "fhnN,hi->finN", "finN,efrR->einNrR", "einNrR,ijsS->ejnNrRsS"],
           "syrnt", "SYRNu", "ejnNrRsS", "ejNRSyt", "ejytYu"),
}


def _half(key, env_ops, T):
    """(pre, mid, full): the half before the ket tensor, after it, after
    the bra tensor."""
    steps, lk, lb, lpre, lmid, lfull = _HALVES[key]
    t = env_ops[0]
    for spec, o in zip(steps, env_ops[1:]):
        t = _es(spec, t, o, optimize=True)
    pre = t
    mid = _es(f"{lpre},{lk}->{lmid}", pre, T, optimize=True)
    full = _es(f"{lmid},{lb}->{lfull}", mid, T, optimize=True)
    return pre, mid, full


def _half_grad(key, pre, mid, T, M):
    """d <full, M> / dT, both occurrences (ket and bra)."""
    _, lk, lb, lpre, lmid, lfull = _HALVES[key]
    return (_es(f"{lpre},{lb},{lfull}->{lk}", pre, T, M, optimize=True)
            + _es(f"{lmid},{lfull}->{lb}", mid, M, optimize=True))


def _bond(env, Tn, lat, source, dx, dy, op):
    """(half keys, env operands of each half, sites, op in (ket1, ket2, bra1,
    bra2) order of the first/second site)."""
    nrow, ncol, _, _, _, _, idx = cluster(lat, source, dx, dy)
    if nrow * ncol != 2:
        raise ValueError("nearest-neighbour bonds only")
    if nrow == 2:
        s1, s2 = idx[0][0], idx[1][0]
        o = op if s1 == source else op.transpose(1, 0, 3, 2)
        e1 = [env.C1[s1], env.eTt[s1], env.C2[s1], env.eTr[s1], env.eTl[s1]]
        e2 = [env.C3[s2], env.eTb[s2], env.C4[s2], env.eTr[s2], env.eTl[s2]]
        return ("v1", "v2"), (e1, e2), (s1, s2), o
    s1, s2 = idx[0][0], idx[0][1]
    o = op if s1 == source else op.transpose(1, 0, 3, 2)
    e1 = [env.C1[s1], env.eTl[s1], env.C4[s1], env.eTb[s1], env.eTt[s1]]
    e2 = [env.C2[s2], env.eTr[s2], env.C3[s2], env.eTb[s2], env.eTt[s2]]
    return ("h1", "h2"), (e1, e2), (s1, s2), o


def reduced_density_matrix(env, Tn, lat, source, dx, dy):
    """Unnormalised two-site density matrix rho[p, q, t, u] (ket1, bra1,
    ket2, bra2) of the bond from `source` to its (dx, dy) neighbour."""
    d = Tn[source].shape[4]
    keys, envs, (s1, s2), _ = _bond(env, Tn, lat, source, dx, dy, np.zeros((d,) * 4))
    _, _, f1 = _half(keys[0], envs[0], Tn[s1])
    _, _, f2 = _half(keys[1], envs[1], Tn[s2])
    return _es("abcpdq,abctdu->pqtu", f1, f2, optimize=True)


def bond_value(env, Tn, lat, source, dx, dy, op):
    """Unnormalised op12 value of one bond (equals
    Contract_two_sites_*_op12_iTPS_CTM with op12 = op)."""
    keys, envs, (s1, s2), o = _bond(env, Tn, lat, source, dx, dy, op)
    _, _, f1 = _half(keys[0], envs[0], Tn[s1])
    _, _, f2 = _half(keys[1], envs[1], Tn[s2])
    rho = _es("abcpdq,abctdu->pqtu", f1, f2, optimize=True)
    return float(_es("pqtu,ptqu->", rho, o))


def energy_and_grad(env, Tn, lat, bonds, want_grad: bool = True):
    """E = (1/N_sites) sum_b <op_b>, <op_b> = num_b / norm_b, and dE/dTn_i.

    bonds: [(source, dx, dy, op12 (d, d, d, d))].  The norm of a bond is the
    same network with the identity operator.  dE_b = (d num_b - E_b d norm_b)
    / norm_b, i.e. one network with op_eff = (op - E_b 1) / norm_b."""
    n = lat.N_UNIT
    grads = [np.zeros_like(T) for T in Tn] if want_grad else None
    E = 0.0
    for source, dx, dy, op in bonds:
        keys, envs, (s1, s2), o = _bond(env, Tn, lat, source, dx, dy, op)
        p1, m1, f1 = _half(keys[0], envs[0], Tn[s1])
        p2, m2, f2 = _half(keys[1], envs[1], Tn[s2])
        rho = _es("abcpdq,abctdu->pqtu", f1, f2, optimize=True)
        eye = np.einsum("ac,bd->abcd", np.eye(o.shape[0]), np.eye(o.shape[1]))
        num = float(_es("pqtu,ptqu->", rho, o))
        den = float(_es("pqtu,ptqu->", rho, eye))
        Eb = num / den
        E += Eb
        if want_grad:
            oe = (o - Eb * eye) / den
            M1 = _es("abctdu,ptqu->abcpdq", f2, oe, optimize=True)
            M2 = _es("abcpdq,ptqu->abctdu", f1, oe, optimize=True)
            grads[s1] += _half_grad(keys[0], p1, m1, Tn[s1], M1)
            grads[s2] += _half_grad(keys[1], p2, m2, Tn[s2], M2)
    E /= n
    if want_grad:
        grads = [g / n for g in grads]
    return E, grads
