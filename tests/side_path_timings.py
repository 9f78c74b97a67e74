"""Wall-clock timings of the paths beside the CTM sweep, GPU against the
numpy oracle on the same inputs (the bench of the sweep itself is bench.py):

  * full-update bond (Full_update_bond, D=5, chi=100 environment from 4 CTM
    sweeps of the bench's Heisenberg tensors), one bond;
  * correlation functions along +x, r = 1..10, D=5 chi=100 (one left site,
    two right operators);
  * energy and gradient of the bond Hamiltonian with the environment fixed,
    D=4 chi=64 random tensors;
  * the finite-temperature pipelines of TeNeS's FT_TFI_square and FT_Kitaev
    inputs (TPO simple update, single-layer CTM and all measurements at every
    measured temperature), end to end.

    python tests/side_path_timings.py

prints one JSON line.  The CPU column is the numpy oracle (the tests'
checker), timed once; the GPU column is the best of three calls.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np


KINDS = ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl")


def _best(fn, reps=3):
    best = float("inf")
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        best = min(best, time.perf_counter() - t0)
    return best, out


def main():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    from tenes_amd import core, inputs
    from oracle import adjoint as A
    from oracle import contract as K
    from oracle import ctm as O
    from oracle import full_update as FU
    from oracle.lattice import SquareLattice

    fx = json.load(open(os.path.join(root, "tests", "golden", "heisenberg_d5_chi100.json")))
    D = fx["D"]
    Tn = [np.array(t).reshape(D, D, D, D, 2) for t in fx["site_tensors"]]
    ctx = core.Context(0, seed=3)
    env = core.Environment(ctx, 2, 2)
    env.set_site_tensors(Tn)
    env.calc_ctm_environment(core.peps_parameters(CHI=100, Max_CTM_Iteration=4, CTM_Convergence_Epsilon=0.0))
    host = O.Env(4)
    for k in KINDS:
        for i in range(4):
            getattr(host, k)[i] = env.get(k, i).numpy()
    out = {}

    # full-update bond, site 0 -> right neighbour (leg 2)
    H = inputs.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    gate = inputs.gate(H, 0.01)
    lat = SquareLattice(2, 2)
    s, t = 0, lat.right(0)
    names = ("C1", "C2", "C3", "C4", "eTt", "eTt", "eTr", "eTb", "eTb", "eTl")
    sites = (s, t, t, s, s, t, t, t, s, s)
    dev = [env.get(k, i) for k, i in zip(names, sites)] + [env.get("Tn", s), env.get("Tn", t)]
    hst = [getattr(host, k)[i] for k, i in zip(names, sites)] + [Tn[s], Tn[t]]

    def gpu_fu():
        a, b = core.full_update_bond(dev, gate, 2)
        ctx.synchronize()
        return a, b

    tg, _ = _best(gpu_fu)
    tc, _ = _best(lambda: FU.Full_update_bond(*hst, gate, 2, FU.FUParameters()), reps=1)
    out["full_update_bond_d5_chi100"] = {"gpu_s": tg, "cpu_oracle_s": tc}

    # correlation functions along +x, r = 1..10
    from tenes_amd.inputs import SX, SZ

    tg, _ = _best(lambda: env.correlation(0, SZ, [SZ, SX], 10, 0))

    def cpu_cor():
        A0 = K.StartCorrelation_iTPS_CTM(host.C1[0], host.C4[0], host.eTt[0], host.eTb[0], host.eTl[0], Tn[0], SZ)
        q = 0
        vals = []
        for _ in range(10):
            q = lat.right(q)
            vals.append(K.FinishCorrelation_iTPS_CTM(A0, host.C2[q], host.C3[q], host.eTt[q], host.eTr[q],
                                                     host.eTb[q], Tn[q], SZ))
            A0 = K.Transfer_iTPS_CTM(A0, host.eTt[q], host.eTb[q], Tn[q])
        return vals

    tc, _ = _best(cpu_cor, reps=1)
    out["correlation_r10_d5_chi100"] = {"gpu_s": tg, "cpu_oracle_s_one_operator_no_norm": tc}

    # energy + gradient at D=4, chi=64
    T4 = inputs.random_tensors(4, 4, 2, seed=11)
    e4 = core.Environment(ctx, 2, 2)
    e4.set_site_tensors(T4)
    e4.calc_ctm_environment(core.peps_parameters(CHI=64, Max_CTM_Iteration=10, CTM_Convergence_Epsilon=0.0))
    h4 = O.Env(4)
    for k in KINDS:
        for i in range(4):
            getattr(h4, k)[i] = e4.get(k, i).numpy()
    bonds = [(i, 1, 0, H) for i in range(4)] + [(i, 0, 1, H) for i in range(4)]
    tg, _ = _best(lambda: e4.bond_energy(bonds, grad=True))
    tc, _ = _best(lambda: A.energy_and_grad(h4, T4, lat, bonds), reps=1)
    out["energy_gradient_d4_chi64"] = {"gpu_s": tg, "cpu_oracle_s": tc}

    # finite temperature, end to end
    sys.path.insert(0, os.path.join(root, "tests"))
    import ft_cases

    for name in ("FT_TFI_square", "FT_Kitaev"):
        case = ft_cases.setup(name)
        t0 = time.perf_counter()
        ft_cases.gpu_records(ctx, case)
        tg = time.perf_counter() - t0
        t0 = time.perf_counter()
        ft_cases.oracle_records(case)
        tc = time.perf_counter() - t0
        out[f"finite_temperature_{name}"] = {"gpu_s": tg, "cpu_oracle_s": tc}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
