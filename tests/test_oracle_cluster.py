"""The oracle's cluster contraction for every shape up to 4x4
(contract_itps_ctm/ctm.cpp:45-87, ctm_NxM.cpp), pinned three ways:

* against the oracle's 1x1, 1x2, 2x1 and 2x2 trees (themselves pinned to
  TeNeS's onesite/twosite/multisite outputs) on random environments;
* at D = 1, where the network factorises into the trace of the clockwise ring
  of corner and edge matrices times one scalar <t|op|t> per site: a closed
  form that fixes every leg order of the ring and every site/operator
  placement;
* against the correlation functions (StartCorrelation / Transfer /
  FinishCorrelation, correlation.cpp:34-105, pinned to TeNeS's
  correlation.dat): a 1 x (r+1) row or (r+1) x 1 column with operators at its
  ends is the same network.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import contract as K
from oracle.ctm import Env
from oracle.itps import _cluster_args
from oracle.lattice import SquareLattice


def random_env(lat, chi, D, d, seed):
    rng = np.random.default_rng(seed)
    N = lat.N_UNIT
    env = Env(N)
    for i in range(N):
        for k in ("C1", "C2", "C3", "C4"):
            getattr(env, k)[i] = rng.standard_normal((chi, chi))
        for k in ("eTt", "eTr", "eTb", "eTl"):
            getattr(env, k)[i] = rng.standard_normal((chi, chi, D, D))
    Tn = [rng.standard_normal((D, D, D, D, d)) for _ in range(N)]
    ops = [rng.standard_normal((d, d)) for _ in range(N)]  # not symmetric
    return env, Tn, ops


def cluster(env, Tn, ops, lat, src, nrow, ncol, with_ops=True):
    """Cluster whose bottom-left site is src (rows top to bottom)."""
    idx = [[lat.other(src, c, nrow - 1 - r) for c in range(ncol)] for r in range(nrow)]
    C, eTt, eTr, eTb, eTl = _cluster_args(env, idx, nrow, ncol)
    T = [[Tn[idx[r][c]] for c in range(ncol)] for r in range(nrow)]
    d = Tn[0].shape[4]
    O = [[(ops[idx[r][c]] if with_ops else np.eye(d)) for c in range(ncol)] for r in range(nrow)]
    return C, eTt, eTr, eTb, eTl, T, O


@pytest.mark.parametrize("shape", [(1, 1), (1, 2), (2, 1), (2, 2)])
def test_sweep_matches_trees(shape):
    lat = SquareLattice(2, 2)
    env, Tn, ops = random_env(lat, 5, 2, 2, seed=1)
    for src in range(lat.N_UNIT):
        args = cluster(env, Tn, ops, lat, src, *shape)
        ref = K.Contract_iTPS_CTM(*args)
        val = K._cluster_rows(*args)
        assert abs(val - ref) <= 1e-12 * abs(ref), (shape, src, val, ref)


@pytest.mark.parametrize("shape", [(r, c) for r in range(1, 5) for c in range(1, 5)])
def test_d1_ring_closed_form(shape):
    nrow, ncol = shape
    lat = SquareLattice(2, 2)
    env, Tn, ops = random_env(lat, 3, 1, 2, seed=7)
    src = 1
    args = cluster(env, Tn, ops, lat, src, nrow, ncol)
    C, eTt, eTr, eTb, eTl, T, O = args
    M = C[0]
    for c in range(ncol):
        M = M @ eTt[c][:, :, 0, 0]
    M = M @ C[1]
    for r in range(nrow):
        M = M @ eTr[r][:, :, 0, 0]
    M = M @ C[2]
    for c in reversed(range(ncol)):
        M = M @ eTb[c][:, :, 0, 0]
    M = M @ C[3]
    for r in reversed(range(nrow)):
        M = M @ eTl[r][:, :, 0, 0]
    expect = np.trace(M)
    for r in range(nrow):
        for c in range(ncol):
            t = T[r][c].reshape(-1)
            expect *= t @ O[r][c] @ t  # op[ket, bra]
    val = K.Contract_iTPS_CTM(*args)
    assert abs(val - expect) <= 1e-12 * abs(expect), (shape, val, expect)


@pytest.mark.parametrize("r", [1, 2, 3])
def test_row_and_column_match_correlation_functions(r):
    lat = SquareLattice(2, 2)
    env, Tn, ops = random_env(lat, 4, 2, 2, seed=3 + r)
    A_op, B_op = ops[0], ops[1]
    rot = (3, 0, 1, 2, 4)
    for left in range(lat.N_UNIT):
        # horizontal: 1 x (r+1), A at the left end, B at the right end
        A = K.StartCorrelation_iTPS_CTM(env.C1[left], env.C4[left], env.eTt[left], env.eTb[left],
                                        env.eTl[left], Tn[left], A_op)
        s = left
        for _ in range(r - 1):
            s = lat.right(s)
            A = K.Transfer_iTPS_CTM(A, env.eTt[s], env.eTb[s], Tn[s])
        s = lat.right(s)
        ref = K.FinishCorrelation_iTPS_CTM(A, env.C2[s], env.C3[s], env.eTt[s], env.eTr[s],
                                           env.eTb[s], Tn[s], B_op)
        C, eTt, eTr, eTb, eTl, T, O = cluster(env, Tn, ops, lat, left, 1, r + 1, with_ops=False)
        O[0][0], O[0][r] = A_op, B_op
        val = K._cluster_rows(C, eTt, eTr, eTb, eTl, T, O)
        assert abs(val - ref) <= 1e-11 * abs(ref), ("row", left, val, ref)
        # vertical: (r+1) x 1, A at the bottom, B at the top
        t = Tn[left].transpose(rot)
        A = K.StartCorrelation_iTPS_CTM(env.C4[left], env.C3[left], env.eTl[left], env.eTr[left],
                                        env.eTb[left], t, A_op)
        s = left
        for _ in range(r - 1):
            s = lat.top(s)
            A = K.Transfer_iTPS_CTM(A, env.eTl[s], env.eTr[s], Tn[s].transpose(rot))
        s = lat.top(s)
        ref = K.FinishCorrelation_iTPS_CTM(A, env.C1[s], env.C2[s], env.eTl[s], env.eTt[s],
                                           env.eTr[s], Tn[s].transpose(rot), B_op)
        C, eTt, eTr, eTb, eTl, T, O = cluster(env, Tn, ops, lat, left, r + 1, 1, with_ops=False)
        O[r][0], O[0][0] = A_op, B_op
        val = K._cluster_rows(C, eTt, eTr, eTb, eTl, T, O)
        assert abs(val - ref) <= 1e-11 * abs(ref), ("column", left, val, ref)


def test_too_large_cluster_is_an_error():
    lat = SquareLattice(2, 2)
    env, Tn, ops = random_env(lat, 2, 1, 2, seed=0)
    with pytest.raises(NotImplementedError):
        K.Contract_iTPS_CTM(*cluster(env, Tn, ops, lat, 0, 1, 5))
