"""Finite-temperature (iTPO) pipeline on the GPU: the TPO simple update
(tenes_env_simple_update on (D,D,D,D,d,d) site tensors), the single-layer
CTM (tenes_calc_ctm_environment_density) and the density contractions,
against TeNeS's outputs for FT_Kitaev and FT_TFI_square (tolerances and
their reasons in ft_cases.tolerance) and against the oracle on the same
TPO tensors."""
import numpy as np
import pytest

import ft_cases as F
from ctm_cases import spectra_diff
import tenes_amd as ta

pytestmark = pytest.mark.gpu

KINDS = ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl")


@pytest.mark.parametrize("name", ["FT_Kitaev", "FT_TFI_square"])
def test_gpu_finite_temperature_goldens(gpu_ctx, name):
    case = F.setup(name)
    recs = F.gpu_records(gpu_ctx, case)
    assert len(recs) == case["nsteps"] // case["interval"] + 1
    diff = F.differences(case, recs)
    bad = {k: v for k, v in diff.items() if v > F.tolerance(name, k[0], k[1], gpu=True)}
    assert not bad, "; ".join(f"{k}: {v:.2e}" for k, v in sorted(diff.items()))


def _oracle_tpo(case, steps):
    from oracle import finite_temperature as FT

    lat = case["lat"]
    Tn = FT.initial_tensors_density(lat, case["pdims"], case["vdims"])
    lam = [[np.ones(v[l]) for l in range(4)] for v in case["vdims"]]
    for _ in range(steps):
        for s, l, op in case["evolution"]:
            FT.simple_update_density(Tn, lam, lat, s, l, op)
    return Tn, lam


def test_gpu_tpo_simple_update_vs_oracle(gpu_ctx):
    """FT_TFI_square, 100 simple-update steps of the TPO (beta = 1, before
    the bond-spectrum crossing): the GPU's bond spectra equal the oracle's
    to 1e-12."""
    from tenes_amd import finite_temperature as FTG

    case = F.setup("FT_TFI_square")
    Tn, lam = _oracle_tpo(case, 100)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(FTG.initial_tpo(case["pdims"], case["vdims"]))
    genv.simple_update(case["evolution"], 100)
    for i in range(4):
        for l in range(4):
            assert np.max(np.abs(genv.get_lambda(i, l) - lam[i][l])) < 1e-12, (i, l)


def test_gpu_single_layer_ctm_vs_oracle(gpu_ctx):
    """FT_TFI_square TPO tensors after 300 oracle simple-update steps
    (beta = 3): the GPU single-layer CTM (same sweep count), its corner
    spectra, the observables and the correlation functions agree with the
    oracle's on the same tensors to 1e-10."""
    from oracle import ctm, ctm_single as CS, itps, contract_density as KD, transfer as TMo
    from tenes_amd import finite_temperature as FTG

    case = F.setup("FT_TFI_square")
    lat = case["lat"]
    Tn, _ = _oracle_tpo(case, 300)
    p = ctm.PEPSParameters(CHI=10, Max_CTM_Iteration=10)
    env, count, _ = CS.Calc_CTM_Environment_density(Tn, p, lat)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    params = ta.peps_parameters(CHI=10, Max_CTM_Iteration=10)
    rec = FTG.measure(genv, 2, 2, 0, case["onesite"], case["twosite"], case["site_ops"], 3, case["pairs"],
                      params, 3.0)
    assert rec["ctm_count"] == count
    for i in range(4):
        for k in KINDS[:4]:
            a = np.linalg.svd(getattr(env, k)[i], compute_uv=False)
            b = np.linalg.svd(genv.get(k, i).numpy(), compute_uv=False)
            assert spectra_diff(a / np.linalg.norm(a), b / np.linalg.norm(b)) < 1e-12, (k, i)
        # edges: the single-layer shape (chi, chi, D) plus a unit bra leg (the
        # chi legs at least the reference's: CHI-sized with the cut on the device)
        ge, oe = genv.get("eTt", i).shape, env.eTt[i].shape
        assert ge[2] == oe[2] and ge[3] == 1 and ge[0] >= oe[0] and ge[1] >= oe[1]
    one, _ = itps.measure_onesite(env, Tn, lat, case["onesite"], K=KD)
    two, _ = itps.measure_twosite(env, Tn, lat, case["twosite"], case["onesite"], K=KD)
    cor = itps.measure_correlation(env, Tn, lat, case["site_ops"], 3, case["pairs"], K=KD)
    for g in one:
        for i in one[g]:
            assert abs(rec["onesite"][g][i] - one[g][i]) < 1e-10, (g, i)
    for g in two:
        for k in two[g]:
            assert abs(rec["twosite"][g][k] - two[g][k]) < 1e-10, (g, k)
    for k in cor:
        assert abs(rec["correlation"][k] - cor[k]) < 1e-10, k
    xi = itps.correlation_lengths(env, Tn, lat, TMo.TMParameters())
    mine = {(x[0], x[1]): x[2] for x in rec["correlation_length"]}
    for d, f, x, _, _ in xi:
        assert abs(mine[(d, f)] - x) < 1e-9 * abs(x), (d, f, mine[(d, f)], x)
    # multisite products of the density network (Contract_density_CTM,
    # contract_density_ctm/ctm.cpp:45-85): a 2x2 plaquette and a 3x3 cluster
    ms = [{"group": 0, "source": 0, "dx": [1, 1, 0], "dy": [0, 1, 1], "ops": [0, 1, 0, 1], "coeff": 1.0},
          {"group": 1, "source": 1, "dx": [2, -1], "dy": [1, -1], "ops": [0, 0, 1], "coeff": 1.0}]
    ref = itps.measure_multisite(env, Tn, lat, case["site_ops"], ms, K=KD)
    for m in ms:
        ops = [case["site_ops"][(lat.other(m["source"], *(([0] + m["dx"])[n], ([0] + m["dy"])[n])), g)]
               for n, g in enumerate(m["ops"])]
        got = genv.contract_multisite(m["source"], m["dx"], m["dy"], [np.real(o) for o in ops])
        key = (m["group"], m["source"], *[v for pair in zip(m["dx"], m["dy"]) for v in pair])
        assert abs(got - ref[key].real) < 1e-10, (key, got, ref[key])


@pytest.mark.parametrize("steps", [100, 300])
@pytest.mark.parametrize("device_cut", [0, 1], ids=["host-cut", "device-cut"])
def test_gpu_single_layer_updown_vs_oracle(gpu_ctx, steps, device_cut):
    """Up-down projectors of the single-layer network (CTM_Projector_corner =
    0, Calc_projector_updown_blocks_single, ctm_single.cpp:267-436, with its
    truncation loop over the first e78 singular values): FT_TFI_square TPO
    tensors after 100 and 300 oracle simple-update steps (beta = 1, 3), CHI
    = 10, run to the convergence test (epsilon 1e-10) on the GPU and in the
    oracle: the same sweep count, corner spectra to 1e-10 (measured 2e-11 at
    beta = 1: the runs stop at a change of 1e-10) and every one-site and
    two-site observable to 1e-10.  With the cut on the device the
    projectors keep CHI-sized shapes with zero columns where the host cut
    shrinks them (the spectra are compared with zero padding)."""
    from oracle import ctm, ctm_single as CS, itps, contract_density as KD
    from tenes_amd import finite_temperature as FTG

    case = F.setup("FT_TFI_square")
    lat = case["lat"]
    Tn, _ = _oracle_tpo(case, steps)
    p = ctm.PEPSParameters(CHI=10, Max_CTM_Iteration=100, CTM_Convergence_Epsilon=1e-10,
                           CTM_Projector_corner=False)
    env, count, _ = CS.Calc_CTM_Environment_density(Tn, p, lat)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    params = ta.peps_parameters(CHI=10, Max_CTM_Iteration=100, CTM_Convergence_Epsilon=1e-10,
                                CTM_Projector_corner=0, projector_cut_on_device=device_cut)
    rec = FTG.measure(genv, 2, 2, 0, case["onesite"], case["twosite"], case["site_ops"], 0, [],
                      params, 1.0)
    assert rec["ctm_count"] == count
    for i in range(4):
        for k in KINDS[:4]:
            a = np.linalg.svd(getattr(env, k)[i], compute_uv=False)
            b = np.linalg.svd(genv.get(k, i).numpy(), compute_uv=False)
            assert spectra_diff(a / np.linalg.norm(a), b / np.linalg.norm(b)) < 1e-10, (k, i)
    one, _ = itps.measure_onesite(env, Tn, lat, case["onesite"], K=KD)
    two, _ = itps.measure_twosite(env, Tn, lat, case["twosite"], case["onesite"], K=KD)
    for g in one:
        for i in one[g]:
            assert abs(rec["onesite"][g][i] - one[g][i]) < 1e-10, (g, i)
    for g in two:
        for k in two[g]:
            assert abs(rec["twosite"][g][k] - two[g][k]) < 1e-10, (g, k)


def test_gpu_error_inside_a_forked_move_restores_the_context():
    """An error raised inside a move whose rows run on side streams (here
    the refusal of a projector SVD beyond 240 triplets: CHI = 250 with
    projector matrices 256 wide from the first sweep, random D=4 tensors):
    the context is back on its own stream afterwards (no fork left open) and
    the environment stays usable.  A context of its own, so that a failure
    here cannot leak into the other tests' session context."""
    from tenes_amd import inputs

    ctx = ta.Context(0, seed=99)
    genv = ta.Environment(ctx, 2, 2)
    genv.set_site_tensors(inputs.random_tensors(4, 4, 2, seed=3))
    assert ctx.lane_state() == (False, -1)
    with pytest.raises(ta.TenesError, match="error -5: .*range-finder block"):
        genv.calc_ctm_environment(ta.peps_parameters(CHI=250, Max_CTM_Iteration=2))
    assert ctx.lane_state() == (False, -1)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=20, Max_CTM_Iteration=2))
    assert ctx.lane_state() == (False, -1)
    assert n >= 1 and np.isfinite(genv.contract_one_site(0))
