"""Batched CTMRG (tenes_batch_calc_ctm_environment): independent problems on
several contexts, run concurrently, against the same problems run one by one
and against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import tenes_amd as ta
from oracle import ctm as O
from oracle.lattice import SquareLattice
from tenes_amd import batch, inputs

pytestmark = pytest.mark.gpu


def _oracle_energy(Tn, H, chi, sweeps):
    from oracle import itps

    lat = SquareLattice(2, 2)
    env, _, _ = O.Calc_CTM_Environment(list(Tn), O.PEPSParameters(CHI=chi, Max_CTM_Iteration=sweeps,
                                                                  CTM_Convergence_Epsilon=0.0), lat)
    two = [{"name": "H", "group": 0, "source": s, "dx": 1 if leg == 2 else 0, "dy": 1 if leg == 1 else 0, "op": H}
           for s, leg in inputs.square_bonds(2, 2)]
    w, _ = itps.measure_twosite(env, Tn, lat, two, [])
    return float(np.real(sum(w[0].values()))) / 4.0


def test_batch_matches_one_by_one_and_oracle():
    """Six TFI problems (hx on [0, 3], (D, chi) cycling (2, 8), (3, 12)),
    40 GPU simple-update steps, 6 CTM sweeps: batched over three contexts,
    one by one on a fourth, and the oracle on the same site tensors.  These
    projector matrices are <= 128 wide (exact Jacobi, no random start), so
    batched and one-by-one agree to rounding; the energies agree with the
    oracle to 1e-10."""
    ctxs = [ta.Context(0, seed=7 + i) for i in range(4)]
    n, sweeps = 6, 6
    envs, hams = batch.prepare(ctxs[:3], n, su_steps=40, shapes=((2, 8), (3, 12)))
    batch.run_ctm(envs, sweeps, True)
    got = [batch.energy(e, H) for e, H in zip(envs, hams)]
    for k, (e, H) in enumerate(zip(envs, hams)):
        one = ta.Environment(ctxs[3], 2, 2)
        Tn = e.site_tensors()
        one.set_site_tensors(Tn)
        one.calc_ctm_environment(ta.peps_parameters(CHI=e.problem[2], Max_CTM_Iteration=sweeps,
                                                    CTM_Convergence_Epsilon=0.0))
        ref1 = batch.energy(one, H)
        assert abs(got[k] - ref1) < 1e-13, (k, got[k], ref1)
        ref = _oracle_energy(Tn, H, e.problem[2], sweeps)
        assert abs(got[k] - ref) < 1e-10 * max(1.0, abs(ref)), (k, e.problem, got[k], ref)


def test_batch_randomized_path_and_errors():
    """Two D=4, chi=48 problems (768-wide projector matrices: randomized range
    finder) batched on two contexts against one by one (different random
    starts: 1e-10), and the error path: a problem without site tensors makes
    the call fail with the library's message."""
    ctxs = [ta.Context(0, seed=31 + i) for i in range(3)]
    envs, hams = batch.prepare(ctxs[:2], 2, su_steps=20, shapes=((4, 48),))
    batch.run_ctm(envs, 3, True)
    for e, H in zip(envs, hams):
        one = ta.Environment(ctxs[2], 2, 2)
        one.set_site_tensors(e.site_tensors())
        one.calc_ctm_environment(ta.peps_parameters(CHI=48, Max_CTM_Iteration=3, CTM_Convergence_Epsilon=0.0))
        assert abs(batch.energy(e, H) - batch.energy(one, H)) < 1e-10
    bad = ta.Environment(ctxs[2], 2, 2)
    with pytest.raises(ta.TenesError, match="site tensor not set"):
        ta.core.batch_calc_ctm_environment([envs[0], bad], ta.peps_parameters(CHI=48, Max_CTM_Iteration=1))
