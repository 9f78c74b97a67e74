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


def _batch48_fixture():
    import json
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "tfi_batch48.json")
    if not os.path.exists(path):
        pytest.skip("fixture missing")
    return json.load(open(path))


def _batch48_envs(fx):
    """The 48 problems' site tensors from the GPU simple update (four
    contexts), then moved to environments of one fresh context."""
    n, shapes = fx["problems"], tuple(tuple(s) for s in fx["shapes"])
    su_ctxs = [ta.Context(0, seed=61 + i) for i in range(4)]
    envs, hams = batch.prepare(su_ctxs, n, su_steps=fx["su_steps"], tau=fx["tau"], seed=fx["seed"],
                               noise=fx["noise"], shapes=shapes)
    ctx = ta.Context(0, seed=71)
    out = []
    for e in envs:
        g = ta.Environment(ctx, 2, 2)
        g.set_site_tensors(e.site_tensors())
        g.problem = e.problem
        out.append(g)
    return ctx, out, hams


def test_tfi_batch48_workload_graph_sweeps():
    """The batched field sweep at its workload: 48 transverse-field Ising
    problems (Jz = -1, hx_k = 3k/47, (D, chi) cycling (3, 32), (4, 48),
    (5, 64)), 300 GPU simple-update steps from the seeded ferro state, then
    the fixture's sweep count on one context in lockstep with captured
    sweeps (every problem captured once, replayed on the context's lanes).
    Each problem's energy per site against the oracle's (numpy/LAPACK, its
    own simple update and exact-SVD CTM: tests/golden/tfi_batch48.json) to
    1e-10 relative; then the same batch with the contractions on the FP32
    MFMA (contraction_fp32 = 1, projectors FP64) and on the bf16 MFMA with
    split hi/lo operands (contraction_fp32 = 3), their energy errors stated
    and bounded."""
    fx = _batch48_fixture()
    ctx, envs, hams = _batch48_envs(fx)
    sweeps = fx["sweeps"]
    batch.run_ctm(envs, sweeps, True, graph_sweeps=1)
    n = len(envs)
    assert ctx.graph_stats() == (n, n * (sweeps - 1))
    ref = [r["energy_per_site"] for r in fx["results"]]
    got = [batch.energy(e, H) for e, H in zip(envs, hams)]
    rel = [abs(g - r) / abs(r) for g, r in zip(got, ref)]
    worst = int(np.argmax(rel))
    print(f"FP64 batch: max rel energy err {max(rel):.2e} (problem {worst}, {envs[worst].problem})")
    assert max(rel) < 1e-10, (worst, envs[worst].problem, got[worst], ref[worst])
    batch.run_ctm(envs, sweeps, True, graph_sweeps=1, contraction_fp32=1)
    mixed = [batch.energy(e, H) for e, H in zip(envs, hams)]
    relm = [abs(g - r) / abs(r) for g, r in zip(mixed, ref)]
    worst = int(np.argmax(relm))
    print(f"mixed batch (contraction_fp32=1): max rel energy err {max(relm):.2e} "
          f"(problem {worst}, {envs[worst].problem}), median {np.median(relm):.1e}")
    assert max(relm) < 1e-6, (worst, envs[worst].problem, mixed[worst], ref[worst])
    # split precision: every contraction of the moves (R included) on the bf16
    # MFMA as three hi/lo products (~2^-16 per product)
    batch.run_ctm(envs, sweeps, True, graph_sweeps=1, contraction_fp32=3)
    split = [batch.energy(e, H) for e, H in zip(envs, hams)]
    rels = [abs(g - r) / abs(r) for g, r in zip(split, ref)]
    worst = int(np.argmax(rels))
    print(f"split bf16 batch (contraction_fp32=3): max rel energy err {max(rels):.2e} "
          f"(problem {worst}, {envs[worst].problem}), median {np.median(rels):.1e}")
    assert max(rels) < 1e-4, (worst, envs[worst].problem, split[worst], ref[worst])
