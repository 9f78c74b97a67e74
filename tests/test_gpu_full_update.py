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
from ctm_cases import spectra_diff, su_tensors
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


@pytest.mark.parametrize("name,tol", [("Honeycomb", 1e-8), ("Honeycomb_skew", 1e-5)])
def test_honeycomb_pipeline_on_gpu(gpu_ctx, name, tol):
    """The Honeycomb inputs (virtual dimension 1 on one leg of every site:
    identity projectors; 2x2 cell and skew 2x1 cell) through simple update
    (100 steps), full update (10 steps, fast-full-update moves) and CTM
    (iteration_max 10) on the GPU path: densities, correlation lengths and
    correlation functions against the oracle running the same pipeline, and
    for the 2x2 cell against TeNeS's outputs (1e-8).  On the skew cell the
    fast full update is not covariant under the bond gauge that each bond
    update's final SVD leaves open (environment tensors of the other cell
    rows keep the old gauge until the next sweep), so two correct
    implementations with different SVD sign conventions drift apart: after 10
    steps the GPU and the oracle differ by 1.6e-7 in energy and 1.4e-6 in
    SzSz (bar 1e-5), the oracle and TeNeS by 1.6e-6 and 1.4e-5; each step
    from identical inputs agrees to 1e-15
    (test_honeycomb_skew_full_update_steps)."""
    from oracle import ctm as Oc
    from oracle import itps
    from oracle import simple_update as SU
    from oracle import transfer as TMo

    fx = G.load(name)
    LX, LY, skew, pdims, vdims, dirs, noises = G.unitcell(fx)
    lat = SquareLattice(LX, LY, skew)
    Tn0 = SU.initial_tensors(lat, pdims, vdims, dirs, noises, seed=11, is_real=True)
    c = fx["parameter"]["ctm"]
    prm = ta.peps_parameters(CHI=c["dimension"], Max_CTM_Iteration=c["iteration_max"])
    genv = ta.Environment(gpu_ctx, LX, LY, skew)
    genv.set_site_tensors(Tn0)
    genv.simple_update(G.evolution(fx, True), fx["parameter"]["simple_update"]["num_step"])
    genv.full_update(G.evolution(fx, True, "evolution_full"), fx["parameter"]["full_update"]["num_step"],
                     ctm_params=prm)
    genv.calc_ctm_environment(prm)

    # oracle, same pipeline
    _, _, _, Tn = G.pipeline_site_tensors(name)
    env, _, _ = Oc.Calc_CTM_Environment(Tn, G.ctm_parameters(fx), lat)
    one, onames, two, tnames = G.operators(fx, lat.N_UNIT, True)
    o, _ = itps.measure_onesite(env, Tn, lat, one)
    w, _ = itps.measure_twosite(env, Tn, lat, two, one)
    dens = itps.density(o, w, lat, pdims, onames, tnames)
    gold = G.density_table(fx["outputs"]["density"], timed=False)

    site_op = {(x["site"], x["group"]): x for x in one}
    vals = {}
    for g, nm in onames.items():
        vals[nm] = sum(x["coeff"] * genv.contract_one_site(x["site"], x["op"]) / genv.contract_one_site(x["site"])
                       for x in one if x["group"] == g) / lat.N_UNIT
    for g, nm in tnames.items():
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
        vals[nm] = tot / lat.N_UNIT
    for nm, v in vals.items():
        assert abs(v - dens[nm].real) < tol, (nm, v, dens[nm])
        if name == "Honeycomb":
            assert abs(v - gold[nm].real) < 1e-8, (nm, v, gold[nm])
    # correlation lengths and correlation functions against the oracle
    for d, f, xi, es, ev in itps.correlation_lengths(env, Tn, lat, TMo.TMParameters()):
        gev = genv.transfer_matrix_eigenvalues(d, f)
        L = lat.LX_noskew if d == 0 else lat.LY_noskew
        gxi = -L / np.log(abs(gev[1]) / abs(gev[0]))
        assert abs(gxi - xi) < 10 * tol * max(1.0, abs(xi)), (d, f, gxi, xi)
    site_ops = {(x["site"], x["group"]): x["op"] for x in one}
    pairs = [tuple(x) for x in fx["correlation"]["operators"]]
    cor = itps.measure_correlation(env, Tn, lat, site_ops, fx["correlation"]["r_max"], pairs)
    for left in range(lat.N_UNIT):
        rgs = [p[1] for p in pairs if p[0] == 0]
        for direction in (0, 1):
            got = genv.correlation(left, site_ops[(left, 0)], [site_ops[(left, rg)] for rg in rgs],
                                   fx["correlation"]["r_max"], direction)
            for r in range(fx["correlation"]["r_max"]):
                for j, rg in enumerate(rgs):
                    key = (0, left, rg, r + 1 if direction == 0 else 0, r + 1 if direction == 1 else 0)
                    assert abs(got[r, j] - cor[key].real) < tol, (key, got[r, j], cor[key])


def test_honeycomb_skew_full_update_steps(gpu_ctx):
    """Skew 2x1 cell: every full-update step from identical inputs (the
    oracle's state after 100 simple-update steps and its CTM environment,
    uploaded): the GPU bond update's two-site tensor and the GPU
    fast-full-update moves' corner spectra (fed the oracle's updated tensors)
    agree with the oracle's step to 1e-12."""
    from oracle import ctm as Oc
    from oracle import simple_update as SU

    KINDS = ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl")
    fx = G.load("Honeycomb_skew")
    LX, LY, skew, pdims, vdims, dirs, noises = G.unitcell(fx)
    lat = SquareLattice(LX, LY, skew)
    Tn = SU.initial_tensors(lat, pdims, vdims, dirs, noises, seed=11, is_real=True)
    lam = [[np.ones(vdims[i][l]) for l in range(4)] for i in range(lat.N_UNIT)]
    Tn, lam = SU.simple_update_steps(Tn, lam, lat, G.evolution(fx, True), 100)
    p = G.ctm_parameters(fx)
    env, _, _ = Oc.Calc_CTM_Environment(Tn, p, lat)
    prm = ta.peps_parameters(CHI=p.CHI, Max_CTM_Iteration=p.Max_CTM_Iteration)
    names = {0: ("C4", "C2", "C1", "C3", "eTb", "eTb", "eTl", "eTt", "eTt", "eTr"),
             1: ("C4", "C1", "C2", "C3", "eTl", "eTl", "eTt", "eTr", "eTr", "eTb"),
             2: ("C1", "C2", "C3", "C4", "eTt", "eTt", "eTr", "eTb", "eTb", "eTl"),
             3: ("C2", "C3", "C4", "C1", "eTr", "eTr", "eTb", "eTl", "eTl", "eTt")}
    for src, leg, op in G.evolution(fx, True, "evolution_full"):
        tgt = lat.neighbor(src, leg)
        args = [getattr(env, k)[i] for k, i in zip(names[leg], (src, tgt, tgt, src, src, tgt, tgt, tgt, src, src))]
        r1, r2 = FU.Full_update_bond(*args, Tn[src], Tn[tgt], op, leg, FU.FUParameters())
        g1, g2 = ta.core.full_update_bond([gpu_ctx.tensor(a) for a in args]
                                          + [gpu_ctx.tensor(Tn[src]), gpu_ctx.tensor(Tn[tgt])], op, leg)
        ref, got = _bond_tensor(r1, r2, leg), _bond_tensor(g1.numpy(), g2.numpy(), leg)
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12
        genv = ta.Environment(gpu_ctx, LX, LY, skew)
        genv.set_site_tensors(Tn)
        for k in KINDS:
            for i in range(lat.N_UNIT):
                genv.set(k, i, getattr(env, k)[i])
        env2 = Oc.Env(lat.N_UNIT)
        for k in KINDS:
            setattr(env2, k, list(getattr(env, k)))
        Tn2 = list(Tn)
        FU.full_update_step(env2, Tn2, lat, src, leg, op, p, FU.FUParameters())
        genv.set("Tn", src, Tn2[src])
        genv.set("Tn", tgt, Tn2[tgt])
        sx, tx, sy, ty = src % LX, tgt % LX, src // LX, tgt // LX
        for d, c in {0: [(1, sx), (0, tx)], 1: [(3, sy), (2, ty)], 2: [(0, sx), (1, tx)], 3: [(2, sy), (3, ty)]}[leg]:
            genv.move(d, c, prm)
        for k in ("C1", "C2", "C3", "C4"):
            for i in range(lat.N_UNIT):
                a, b = Oc.normalized_spectrum(getattr(env2, k)[i]), genv.corner_spectrum(k, i)
                assert spectra_diff(a, b) < 1e-12, (src, leg, k, i)
        Tn, env = Tn2, env2


def _gpu_values(genv, lat, fx):
    """Per-group densities and per-bond values measured on the GPU
    environment as measure_onesite / measure_twosite / save_density do
    (twosite_obs.cpp:146-262; explicit operators on longer bonds through the
    SVD split of twosite_obs.cpp:223-243)."""
    one, onames, two, tnames = G.operators(fx, lat.N_UNIT, True)
    site_op = {(o["site"], o["group"]): o for o in one}
    dens, per_site, per_bond = {}, {}, {}
    for g, name in onames.items():
        tot = 0.0
        for o in (x for x in one if x["group"] == g):
            v = o["coeff"] * genv.contract_one_site(o["site"], o["op"]) / genv.contract_one_site(o["site"])
            per_site[(g, o["site"])] = v
            tot += v
        dens[name] = tot / lat.N_UNIT
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
            per_bond[(g, s, dx, dy)] = op["coeff"] * v / norm
            tot += per_bond[(g, s, dx, dy)]
        dens[name] = dens.get(name, 0.0) + tot / lat.N_UNIT
    return dens, per_site, per_bond


def test_j1j2_full_pipeline_on_gpu(gpu_ctx):
    """J1J2_AFH (2x2, D=2, chi=5, J2 = 0.1): 100 simple-update steps whose
    second-neighbour gates are path-decomposed ((2,2,2,8) then (8,2,2,2): the
    intermediate site's physical leg is 8 between them), 1 full-update step
    with the same gates and fast-full-update moves, the CTM and the
    observables, diagonal bonds included, all on the GPU, against TeNeS's
    output_J1J2_AFH.  The full-update steps of the path gates amplify rounding
    (a 1e-15 relative input change moves the energy by ~3e-7 in the oracle,
    tests/test_oracle_goldens.py::test_j1j2_end_to_end_goldens), so the bar
    is 5e-6, as for the oracle."""
    from oracle import simple_update as SU

    fx = G.load("J1J2_AFH")
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
    dens, per_site, per_bond = _gpu_values(genv, lat, fx)
    gold = G.density_table(fx["outputs"]["density"], timed=False)
    # this output names the bond term "hamiltonian" (older TeNeS); it equals Energy
    common = [k for k in dens if k in gold]
    assert len(common) >= 5
    worst = max(abs(dens[k] - gold[k].real) for k in common)
    worst = max(worst, abs(dens["bond_hamiltonian"] - gold["Energy"].real))
    print(f"J1J2 on the GPU: worst density difference against density.dat {worst:.2e}")
    for k in common:
        assert abs(dens[k] - gold[k].real) < 5e-6, (k, dens[k], gold[k])
    assert abs(dens["bond_hamiltonian"] - gold["Energy"].real) < 5e-6
    assert abs(dens["bond_hamiltonian"] - gold["hamiltonian"].real) < 5e-6
    rows = {tuple(int(x) for x in r.split()[:4]): float(r.split()[4]) for r in fx["outputs"]["twosite_obs"]}
    assert sum(abs(dx) + abs(dy) == 2 for (_, _, dx, dy) in per_bond) == 8
    for k, v in per_bond.items():
        assert abs(v - rows[k]) < 5e-6, (k, v, rows[k])


def test_j1j2_path_gates_simple_update_and_observables_match_oracle(gpu_ctx):
    """The well-conditioned parts of the J1J2 pipeline at full precision:
    (1) 100 simple-update steps with the path-decomposed gates on the GPU
    against the oracle's (bond weights, gauge invariant, to 1e-10; the GPU
    QR of a widened site, 8 rows x 16 columns, takes the Q = 1 form);
    (2) the oracle's final site tensors (after its full update) through the
    GPU CTM and observables, diagonal J2 bonds through the SVD split, against
    the oracle on the same tensors to 1e-10."""
    from oracle import itps, simple_update as SU

    fx = G.load("J1J2_AFH")
    LX, LY, skew, pdims, vdims, dirs, noises = G.unitcell(fx)
    lat = SquareLattice(LX, LY, skew)
    T0 = SU.initial_tensors(lat, pdims, vdims, dirs, noises, seed=11, is_real=True)
    lam = [[np.ones(vdims[i][l]) for l in range(4)] for i in range(lat.N_UNIT)]
    ev = G.evolution(fx, True)
    nsu = fx["parameter"]["simple_update"]["num_step"]
    _, lam = SU.simple_update_steps(list(T0), lam, lat, ev, nsu)
    genv = ta.Environment(gpu_ctx, LX, LY, skew)
    genv.set_site_tensors(T0)
    genv.simple_update(ev, nsu)
    worst = 0.0
    for i in range(lat.N_UNIT):
        for leg in range(4):
            worst = max(worst, float(np.max(np.abs(genv.get_lambda(i, leg) - lam[i][leg]))))
    print(f"J1J2 simple update: worst bond-weight difference {worst:.2e}")
    assert worst < 1e-10
    fx, lat, pdims, Tn = G.pipeline_site_tensors("J1J2_AFH")
    chi = fx["parameter"]["ctm"]["dimension"]
    env, _, _ = O.Calc_CTM_Environment(Tn, G.ctm_parameters(fx), lat)
    one, onames, two, tnames = G.operators(fx, lat.N_UNIT, True)
    o, _ = itps.measure_onesite(env, Tn, lat, one)
    w, _ = itps.measure_twosite(env, Tn, lat, two, one)
    ref = itps.density(o, w, lat, pdims, onames, tnames)
    genv = ta.Environment(gpu_ctx, LX, LY, skew)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=chi))
    dens, per_site, per_bond = _gpu_values(genv, lat, fx)
    for name, v in dens.items():
        assert abs(v - ref[name].real) < 1e-10 * max(1.0, abs(ref[name])), (name, v, ref[name])
    for (g, s, dx, dy), v in per_bond.items():
        assert abs(v - w[g][(s, dx, dy)].real) < 1e-10, (g, s, dx, dy, v)
