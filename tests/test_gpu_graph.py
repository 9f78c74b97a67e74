"""Captured CTM sweeps (tenes_peps_parameters.graph_sweeps): the sweep of an
environment issued into a hipGraph once and replayed.  Against the oracle on
the same inputs, against uncaptured sweeps, and the conditions under which a
replay must not happen (a tensor handed out, a tensor replaced)."""
import numpy as np
import pytest

import tenes_amd as ta
from ctm_cases import random_tensors, spectra_diff, su_tensors
from oracle import ctm as O, itps, models
from oracle.lattice import SquareLattice

pytestmark = pytest.mark.gpu


def _energy(genv, ham):
    E = 0.0
    for s, leg in models.square_bonds(SquareLattice(2, 2)):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        E += genv.contract_two_sites_op12(s, dx, dy, ham) / genv.contract_cluster(s, dx, dy)
    return E / 4


def _oracle(Tn, chi, sweeps, ham):
    lat = SquareLattice(2, 2)
    env, _, _ = O.Calc_CTM_Environment(list(Tn), O.PEPSParameters(CHI=chi, Max_CTM_Iteration=sweeps,
                                                                  CTM_Convergence_Epsilon=0.0), lat)
    two = [{"name": "H", "group": 0, "source": s, "dx": 1 if leg == 2 else 0, "dy": 1 if leg == 1 else 0,
            "op": ham} for s, leg in models.square_bonds(lat)]
    w, _ = itps.measure_twosite(env, list(Tn), lat, two, [])
    return env, float(np.real(sum(w[0].values()))) / 4


@pytest.mark.parametrize("D,chi,sweeps,spec_tol", [(3, 20, 6, 1e-10), (4, 40, 5, 1e-7)])
def test_graph_sweeps_match_oracle(D, chi, sweeps, spec_tol):
    """Heisenberg simple-update tensors; projector matrices wider than 128
    (the randomized range finder, warm-started): the first sweep runs
    uncaptured, the second is captured (its launch counts as a replay), the
    rest replay the graph.  Against the same sweeps uncaptured (same seed:
    the same arithmetic, 1e-12) and the oracle: energy to 1e-10, corner
    spectra to spec_tol (D=4, chi=40 after 5 sweeps: the smallest kept
    singular values carry the range finder's 1e-9 to 3e-8)."""
    Tn, ham = su_tensors("heisenberg", D, 60)
    env, E_ref = _oracle(Tn, chi, sweeps, ham)
    out = []
    for graph in (1, 0):
        ctx = ta.Context(0, seed=3)
        genv = ta.Environment(ctx, 2, 2)
        genv.set_site_tensors(list(Tn))
        n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=chi, Max_CTM_Iteration=sweeps,
                                                            CTM_Convergence_Epsilon=0.0, graph_sweeps=graph,
                                                            projector_cut_on_device=1))
        assert n == sweeps
        assert ctx.graph_stats() == ((1, sweeps - 1) if graph else (0, 0))
        out.append((_energy(genv, ham), [genv.corner_spectrum(k, i) for i in range(4)
                                         for k in ("C1", "C2", "C3", "C4")]))
    (E, spec), (E0, spec0) = out
    assert abs(E - E0) < 1e-12 * abs(E0)
    assert max(spectra_diff(a, b) for a, b in zip(spec, spec0)) < 1e-12
    assert abs(E - E_ref) < 1e-10 * abs(E_ref), (E, E_ref)
    ref = [O.normalized_spectrum(getattr(env, k)[i]) for i in range(4) for k in ("C1", "C2", "C3", "C4")]
    assert max(spectra_diff(a, b) for a, b in zip(spec, ref)) < spec_tol


def test_graph_sweeps_continue_across_calls_and_respect_handles():
    """A graph captured in one call is replayed by the next call that
    continues the environment (initialize = 0).  A tensor handed out between
    calls shares the environment's buffer: the next sweep must not overwrite
    it in place, so that sweep runs uncaptured and the handle keeps its
    values; replacing a tensor (tenes_env_set) changes the buffers and the
    sweep is captured again.  Energies against one uncaptured run."""
    Tn, ham = su_tensors("heisenberg", 3, 60)
    kw = dict(CHI=20, CTM_Convergence_Epsilon=0.0)
    ref = ta.Environment(ta.Context(0, seed=5), 2, 2)
    ref.set_site_tensors(list(Tn))
    ref.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=3, projector_cut_on_device=1, **kw))
    ctx = ta.Context(0, seed=5)
    genv = ta.Environment(ctx, 2, 2)
    genv.set_site_tensors(list(Tn))
    genv.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=3, graph_sweeps=1, **kw))
    assert ctx.graph_stats() == (1, 2)
    assert abs(_energy(genv, ham) - _energy(ref, ham)) < 1e-11
    # continue: replays
    genv.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=2, graph_sweeps=1, **kw), initialize=False)
    ref.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=2, projector_cut_on_device=1, **kw),
                             initialize=False)
    assert ctx.graph_stats() == (1, 4)
    assert abs(_energy(genv, ham) - _energy(ref, ham)) < 1e-11
    # a handle to C1[0]: the next sweep runs uncaptured, the handle is intact
    h = genv.get("C1", 0)
    before = h.numpy().copy()
    genv.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=1, graph_sweeps=1, **kw), initialize=False)
    ref.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=1, projector_cut_on_device=1, **kw),
                             initialize=False)
    assert np.array_equal(h.numpy(), before)
    assert ctx.graph_stats() == (1, 4)
    del h
    # new buffers after that uncaptured sweep: captured again, then replayed
    genv.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=2, graph_sweeps=1, **kw), initialize=False)
    ref.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=2, projector_cut_on_device=1, **kw),
                             initialize=False)
    assert ctx.graph_stats() == (2, 6)
    assert abs(_energy(genv, ham) - _energy(ref, ham)) < 1e-11


def test_graph_sweeps_d5_chi100_energy():
    """The benchmark problem (Heisenberg 2x2, D=5, chi=100, the oracle's
    500-step simple-update tensors) with captured sweeps: energy after the
    oracle's sweep count against its FP64 LAPACK run to 1e-10."""
    import json
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "heisenberg_d5_chi100.json")
    if not os.path.exists(path):
        pytest.skip("fixture missing")
    fx = json.load(open(path))
    D = fx["D"]
    Tn = [np.array(t).reshape(D, D, D, D, 2) for t in fx["site_tensors"]]
    ctx = ta.Context(0, seed=11)
    genv = ta.Environment(ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=fx["chi"], Max_CTM_Iteration=fx["sweeps"],
                                                        CTM_Convergence_Epsilon=0.0, graph_sweeps=1))
    assert n == fx["sweeps"]
    captures, replays = ctx.graph_stats()
    assert captures >= 1 and replays >= fx["sweeps"] - 2
    ham = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    E = _energy(genv, ham)
    assert abs(E - fx["energy_per_site"]) < 1e-10 * abs(fx["energy_per_site"]), (E, fx["energy_per_site"])
