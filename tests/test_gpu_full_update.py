"""Full update on the GPU path (tenes_full_update_bond, tenes_env_full_update)
against TeNeS's bond golden (test/data/full_update.dat), the oracle's bond
update on a CTM environment, and TeNeS's AntiferroHeisenberg_real outputs
with every stage (simple update, full update with fast-full-update moves, CTM,
observables) on the GPU."""
from __future__ import annotations

import numpy as np
import pytest

import golden_io as G
import tenes_amd as ta
from ctm_cases import su_tensors
from oracle import ctm as O
from oracle import full_update as FU
from oracle.lattice import SquareLattice

pytestmark = pytest.mark.gpu


def _bond_tensor(T1, T2, leg):
    """Two-site tensor of the pair: contraction over the bond (invariant
    under the bond gauge that the final SVD of the update leaves open)."""
    return np.tensordot(T1, T2, ([leg], [(leg + 2) % 4]))


def test_full_update_bond_golden(gpu_ctx):
    """test/full_update.cpp inputs (D=2, chi=4, corners with ones on the
    leading D x D block, all-ones edges and site tensors, gate op(i,i,i,i) = 1,
    connect = 2): the updated tensors equal the golden values up to a sign
    per (i,j,k,l) block at relative tolerance 1e-8, as the reference test
    compares them."""
    D, d, chi = 2, 2, 4
    C = []
    for _ in range(4):
        c = np.zeros((chi, chi))
        c[:D, :D] = 1.0
        C.append(gpu_ctx.tensor(c))
    E = [gpu_ctx.tensor(np.ones((chi, chi, D, D))) for _ in range(6)]
    T = [gpu_ctx.tensor(np.ones((D, D, D, D, d))) for _ in range(2)]
    op = np.zeros((d, d, d, d))
    for i in range(d):
        op[i, i, i, i] = 1.0
    n1, n2 = ta.core.full_update_bond(C + E + T, op, 2)
    vals = [float(x) for row in G.load("full_update") for x in row]
    ans = [np.array(vals[:32]).reshape(D, D, D, D, d), np.array(vals[32:64]).reshape(D, D, D, D, d)]
    for res, a in zip((n1.numpy(), n2.numpy()), ans):
        for idx in np.ndindex(D, D, D, D):
            r, x = res[idx], a[idx]
            nz = np.nonzero(np.abs(r) > 1e-14)[0]
            sign = 1.0 if nz.size == 0 or r[nz[0]] * x[nz[0]] > 0 else -1.0
            assert np.allclose(sign * r, x, rtol=1e-8, atol=1e-12), (idx, r, x)


@pytest.mark.parametrize("leg", [0, 1, 2, 3])
def test_full_update_bond_vs_oracle(gpu_ctx, leg):
    """Heisenberg D=3 simple-update tensors, the oracle's chi=12 environment
    (uploaded), the Heisenberg gate exp(-0.01 H) on the bond of site 0 along
    `leg`: the updated pair's two-site tensor agrees with the oracle's to
    1e-9 relative (host FP64 eigen/QR/SVD against LAPACK)."""
    from tenes_amd import inputs

    Tn, ham = su_tensors("heisenberg", 3, 50)
    Tn = list(Tn)
    lat = SquareLattice(2, 2)
    env, _, _ = O.Calc_CTM_Environment(Tn, O.PEPSParameters(CHI=12, Max_CTM_Iteration=4,
                                                            CTM_Convergence_Epsilon=0.0), lat)
    gate = inputs.gate(ham, 0.01)
    s, t = 0, lat.neighbor(0, leg)
    e = env
    names = {0: ("C4", "C2", "C1", "C3", "eTb", "eTb", "eTl", "eTt", "eTt", "eTr"),
             1: ("C4", "C1", "C2", "C3", "eTl", "eTl", "eTt", "eTr", "eTr", "eTb"),
             2: ("C1", "C2", "C3", "C4", "eTt", "eTt", "eTr", "eTb", "eTb", "eTl"),
             3: ("C2", "C3", "C4", "C1", "eTr", "eTr", "eTb", "eTl", "eTl", "eTt")}[leg]
    sites = (s, t, t, s, s, t, t, t, s, s)
    args = [getattr(e, k)[i] for k, i in zip(names, sites)]
    r1, r2 = FU.Full_update_bond(*args, Tn[s], Tn[t], gate, leg, FU.FUParameters())
    g1, g2 = ta.core.full_update_bond([gpu_ctx.tensor(a) for a in args] + [gpu_ctx.tensor(Tn[s]),
                                                                          gpu_ctx.tensor(Tn[t])], gate, leg)
    ref = _bond_tensor(r1, r2, leg)
    got = _bond_tensor(g1.numpy(), g2.numpy(), leg)
    assert got.shape == ref.shape
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < 1e-9, err


def test_afh_real_full_pipeline_on_gpu(gpu_ctx):
    """AntiferroHeisenberg_real (2x2, D=2, chi=5): 100 simple-update steps,
    1 full-update step with fast-full-update moves and the CTM all on the
    GPU from TeNeS's seeded initial tensors; energy and the one-site and bond
    densities against TeNeS's own density.dat to 1e-8."""
    from oracle import simple_update as SU

    fx = G.load("AntiferroHeisenberg_real")
    LX, LY, skew, pdims, vdims, dirs, noises = G.unitcell(fx)
    lat = SquareLattice(LX, LY, skew)
    Tn = SU.initial_tensors(lat, pdims, vdims, dirs, noises, seed=11, is_real=True)
    chi = fx["parameter"]["ctm"]["dimension"]
    genv = ta.Environment(gpu_ctx, LX, LY, skew)
    genv.set_site_tensors(Tn)
    genv.simple_update(G.evolution(fx, True), fx["parameter"]["simple_update"]["num_step"])
    genv.full_update(G.evolution(fx, True, "evolution_full"), fx["parameter"]["full_update"]["num_step"],
                     ctm_params=ta.peps_parameters(CHI=chi))
    genv.calc_ctm_environment(ta.peps_parameters(CHI=chi))
    one, onames, two, tnames = G.operators(fx, lat.N_UNIT, True)
    gold = G.density_table(fx["outputs"]["density"], timed=False)
    site_op = {(o["site"], o["group"]): o for o in one}
    vals = {}
    for g, name in onames.items():
        vals[name] = sum(o["coeff"] * genv.contract_one_site(o["site"], o["op"]) / genv.contract_one_site(o["site"])
                         for o in one if o["group"] == g) / 4
    for g, name in tnames.items():
        tot = 0.0
        for op in (t for t in two if t["group"] == g):
            s, dx, dy = op["source"], op["dx"], op["dy"]
            norm = genv.contract_cluster(s, dx, dy)
            if "op" in op:
                v = genv.contract_two_sites_op12(s, dx, dy, op["op"])
            else:
                tgt = lat.other(s, dx, dy)
                v = genv.contract_cluster(s, dx, dy, site_op[(s, op["ops"][0])]["op"],
                                          site_op[(tgt, op["ops"][1])]["op"])
            tot += op["coeff"] * v / norm
        vals[name] = tot / 4
    worst = max(abs(v - gold[name].real) for name, v in vals.items())
    print(f"worst density difference against density.dat: {worst:.2e}")
    for name, v in vals.items():
        assert abs(v - gold[name].real) < 1e-8, (name, v, gold[name])
    assert abs(vals["bond_hamiltonian"] - gold["Energy"].real) < 1e-8
