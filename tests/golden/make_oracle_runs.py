"""Oracle (numpy/LAPACK FP64) reference runs for the larger CTM cases, written
as JSON fixtures next to this script.  These are the build's own oracle
outputs (no reference code runs); the GPU tests and bench.py compare with them.

  ctm_random_d4_chi64.json   2x2 cell, D=4, chi=64, unit-norm random site
      tensors (tenes_amd.inputs.random_tensors, seed 11), environment from
      the site tensors, 30 sweeps with epsilon = 0; L2-normalised singular
      spectra of C1..C4 of every site.

  heisenberg_d5_chi100.json  2x2 Heisenberg AFM, D=5: site tensors from 500
      simple-update steps (tau = 0.01, Neel state + 0.01 noise, seed 11),
      then chi=100 CTM for a fixed number of sweeps (epsilon = 0); energy per
      site, <Sz>, <Sx> per site, the staggered magnetisation, the 10 leading
      transfer-matrix eigenvalues and correlation lengths (Arnoldi, chi^2 =
      10^4) and corner spectra.  The site tensors are stored so that the GPU
      runs the same problem.

  tfi_batch48.json  the batched transverse-field Ising field sweep: 48
      problems, Jz = -1, hx_k = 3k/47, 2x2 cell, (D, chi) cycling (3, 32),
      (4, 48), (5, 64); site tensors from 300 simple-update steps (tau =
      0.01, ferro state along +z + 0.01 noise, seed 11), then chi CTM
      sweeps (a fixed count, epsilon = 0); each problem's energy per site.
      The GPU runs its own simple update from the same initial tensors (the
      bond SVD's sign convention may differ: a gauge change, the energy is
      the same).

Run:  python tests/golden/make_oracle_runs.py [random|heisenberg|batch]  (minutes to
about an hour of CPU; not part of the test suite)."""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import ctm, itps, models, simple_update as SU, transfer as TM  # noqa: E402
from oracle.lattice import SquareLattice  # noqa: E402
from tenes_amd import inputs  # noqa: E402

OUT = Path(__file__).resolve().parent
KINDS = ("C1", "C2", "C3", "C4")


def spectra(env):
    return {f"{k}{i}": list(ctm.normalized_spectrum(getattr(env, k)[i])) for k in KINDS for i in range(4)}


def random_case():
    lat = SquareLattice(2, 2)
    Tn = inputs.random_tensors(4, 4, 2, seed=11)
    p = ctm.PEPSParameters(CHI=64, Max_CTM_Iteration=30, CTM_Convergence_Epsilon=0.0)
    t0 = time.time()
    env, count, _ = ctm.Calc_CTM_Environment(Tn, p, lat)
    print(f"random D=4 chi=64: {count} sweeps in {time.time() - t0:.0f} s", flush=True)
    fx = {"D": 4, "chi": 64, "seed": 11, "sweeps": count, "spectra": spectra(env)}
    (OUT / "ctm_random_d4_chi64.json").write_text(json.dumps(fx))


def heisenberg_case(sweeps: int = 12):
    lat = SquareLattice(2, 2)
    D = 5
    Tn = inputs.initial_tensors(4, D, 2, inputs.neel_dirs(2, 2), 0.01, 11)
    lam = [[np.ones(D) for _ in range(4)] for _ in range(4)]
    H = inputs.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    g = inputs.gate(H, 0.01)
    Tn, lam = SU.simple_update_steps(Tn, lam, lat, [(s, leg, g) for s, leg in inputs.square_bonds(2, 2)], 500)
    p = ctm.PEPSParameters(CHI=100, Max_CTM_Iteration=sweeps, CTM_Convergence_Epsilon=0.0)
    t0 = time.time()
    env, count, _ = ctm.Calc_CTM_Environment(Tn, p, lat)
    print(f"heisenberg D=5 chi=100: {count} sweeps in {time.time() - t0:.0f} s", flush=True)
    one = [{"name": n, "group": gi, "site": i, "op": op}
           for gi, (n, op) in enumerate([("Sz", models.Sz), ("Sx", models.Sx)]) for i in range(4)]
    two = [{"name": "H", "group": 0, "source": s, "dx": 1 if leg == 2 else 0, "dy": 1 if leg == 1 else 0,
            "op": H} for s, leg in inputs.square_bonds(2, 2)]
    o, _ = itps.measure_onesite(env, Tn, lat, one)
    w, _ = itps.measure_twosite(env, Tn, lat, two, one)
    energy = float(np.real(sum(w[0].values()))) / 4.0
    sz = [float(np.real(o[0][i])) for i in range(4)]
    sx = [float(np.real(o[1][i])) for i in range(4)]
    # staggered magnetisation with TeNeS's sublattice sign (SquareLattice parity)
    par = [1, -1, -1, 1]
    ms = float(np.hypot(sum(p_ * v for p_, v in zip(par, sz)), sum(p_ * v for p_, v in zip(par, sx))) / 4.0)
    corr = []
    for d, fixed, xi, es, ev in itps.correlation_lengths(env, Tn, lat, TM.TMParameters(num_eigvals=10)):
        corr.append({"dir": d, "fixed": fixed, "xi": float(xi), "eigvals_re": [float(np.real(x)) for x in ev],
                     "eigvals_im": [float(np.imag(x)) for x in ev]})
    fx = {"D": D, "chi": 100, "su_steps": 500, "tau": 0.01, "seed": 11, "noise": 0.01, "sweeps": count,
          "energy_per_site": energy, "sz": sz, "sx": sx, "staggered_magnetization": ms,
          "correlation": corr, "spectra": spectra(env),
          "site_tensors": [T.ravel().tolist() for T in Tn]}
    (OUT / "heisenberg_d5_chi100.json").write_text(json.dumps(fx))
    print(f"E/site = {energy:.15f}", flush=True)


def _batch_problem(args):
    k, n, shape, su_steps, sweeps = args
    from threadpoolctl import threadpool_limits

    with threadpool_limits(2):
        lat = SquareLattice(2, 2)
        D, chi = shape
        hx = 3.0 * k / (n - 1)
        Tn = inputs.initial_tensors(4, D, 2, [[1.0, 0.0]] * 4, 0.01, 11)
        lam = [[np.ones(D) for _ in range(4)] for _ in range(4)]
        H = inputs.bond_hamiltonian(Jz=-1.0, hx=hx)
        g = inputs.gate(H, 0.01)
        Tn, lam = SU.simple_update_steps(Tn, lam, lat, [(s, leg, g) for s, leg in inputs.square_bonds(2, 2)],
                                         su_steps)
        p = ctm.PEPSParameters(CHI=chi, Max_CTM_Iteration=sweeps, CTM_Convergence_Epsilon=0.0)
        env, count, _ = ctm.Calc_CTM_Environment(Tn, p, lat)
        two = [{"name": "H", "group": 0, "source": s, "dx": 1 if leg == 2 else 0, "dy": 1 if leg == 1 else 0,
                "op": H} for s, leg in inputs.square_bonds(2, 2)]
        w, _ = itps.measure_twosite(env, Tn, lat, two, [])
        energy = float(np.real(sum(w[0].values()))) / 4.0
    print(f"problem {k}: hx={hx:.4f} D={D} chi={chi} E={energy:.15f}", flush=True)
    return {"k": k, "hx": hx, "D": D, "chi": chi, "sweeps": count, "energy_per_site": energy}


def batch_case(n: int = 48, su_steps: int = 300, sweeps: int = 8, workers: int = 4):
    from multiprocessing import Pool

    shapes = ((3, 32), (4, 48), (5, 64))
    jobs = [(k, n, shapes[k % 3], su_steps, sweeps) for k in range(n)]
    # the D=5 problems first: they take longest
    jobs.sort(key=lambda j: -j[2][0])
    t0 = time.time()
    with Pool(workers) as pool:
        out = pool.map(_batch_problem, jobs, chunksize=1)
    out.sort(key=lambda r: r["k"])
    print(f"batch: {n} problems in {time.time() - t0:.0f} s", flush=True)
    fx = {"problems": n, "shapes": [list(s) for s in shapes], "Jz": -1.0, "hx_max": 3.0, "su_steps": su_steps,
          "tau": 0.01, "seed": 11, "noise": 0.01, "sweeps": sweeps, "results": out}
    (OUT / "tfi_batch48.json").write_text(json.dumps(fx, indent=0))


if __name__ == "__main__":
    which = sys.argv[1:] or ["random", "heisenberg"]
    if "batch" in which:
        batch_case()
    if "random" in which:
        random_case()
    if "heisenberg" in which:
        heisenberg_case()
