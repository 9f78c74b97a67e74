"""The oracle's finite-temperature (iTPO) pipeline against TeNeS's outputs
for its two finite-temperature test inputs (test/data/FT_TFI_square.toml,
FT_Kitaev.toml; outputs output_FT_*/FT_*.dat): TPO initial state, simple
update of the density operator, single-layer CTM, one-site and two-site
observables, energy, correlation functions and correlation lengths at every
measured inverse temperature.  Tolerances per table and beta are in
ft_cases.tolerance, with their reasons."""
import pytest

import ft_cases as F


@pytest.mark.parametrize("name", ["FT_Kitaev", "FT_TFI_square"])
def test_oracle_finite_temperature_goldens(name):
    case = F.setup(name)
    recs = F.oracle_records(case)
    betas = sorted({round(r["beta"], 8) for r in recs})
    assert len(betas) == case["nsteps"] // case["interval"] + 1
    diff = F.differences(case, recs)
    tables = {t for _, t in diff}
    assert {"density", "onesite", "twosite", "correlation", "correlation_length"} <= tables
    bad = {k: v for k, v in diff.items() if v > F.tolerance(name, k[0], k[1])}
    assert not bad, bad


def test_single_layer_updown_truncation_known_answers():
    """The truncation loop of the single-layer up-down projectors
    (ctm_single.cpp:396-415) on hand-worked spectra: s_c is resized to cut0
    zeros, the loop runs over the first e78 values, the first value at or
    below the cutoff becomes the cut, and resize then keeps the first cut0
    entries and value-initialises the rest."""
    import numpy as np
    from oracle.ctm_single import _truncate_e78

    def r(x):
        return 1.0 / np.sqrt(x)

    s = np.array([1.0, 0.5, 0.25, 1e-14, 1e-15, 1e-16])
    # small value inside both bounds: the cut stops there
    cut, sc = _truncate_e78(s, 5, 6, 1e-12)
    assert cut == 3 and np.allclose(sc, [1.0, r(0.5), 2.0])
    # e78 < cut0 and nothing small within e78: cut0 entries, zero from e78 on
    s2 = np.array([1.0, 0.5, 0.25, 0.125, 0.0625])
    cut, sc = _truncate_e78(s2, 4, 2, 1e-12)
    assert cut == 4 and np.allclose(sc, [1.0, r(0.5), 0.0, 0.0])
    # the small value lies beyond cut0 (e78 > cut0): the cut grows to it,
    # entries from cut0 on are zero
    s3 = np.array([1.0, 0.5, 0.25, 0.125, 1e-14])
    cut, sc = _truncate_e78(s3, 2, 5, 1e-12)
    assert cut == 4 and np.allclose(sc, [1.0, r(0.5), 0.0, 0.0])
    # nothing small, e78 >= cut0: the double layer's rule
    cut, sc = _truncate_e78(s2, 3, 5, 1e-12)
    assert cut == 3 and np.allclose(sc, [1.0, r(0.5), 2.0])


@pytest.mark.parametrize("steps,tol", [(100, 2e-7), (300, 1e-9)])
def test_single_layer_updown_agrees_with_corner_projectors(steps, tol):
    """The oracle's single-layer CTM with up-down projectors against the
    same CTM with corner projectors (the latter pinned to TeNeS's
    FT_TFI_square outputs above): both converged (epsilon 1e-10, CHI = 10)
    on the FT_TFI_square TPO after 100 and 300 simple-update steps, every
    one-site and two-site observable agrees (measured 6.8e-8 and 5.2e-11:
    the two schemes truncate differently, the fixed point is the same).  A
    wrong leg or operand in the four enlarged corners or the products gives
    another network and fails this."""
    import numpy as np
    from oracle import ctm, ctm_single as CS, itps, contract_density as KD, finite_temperature as FT

    case = F.setup("FT_TFI_square")
    lat = case["lat"]
    Tn = FT.initial_tensors_density(lat, case["pdims"], case["vdims"])
    lam = [[np.ones(v[l]) for l in range(4)] for v in case["vdims"]]
    for _ in range(steps):
        for s, l, op in case["evolution"]:
            FT.simple_update_density(Tn, lam, lat, s, l, op)
    vals = {}
    for corner in (True, False):
        p = ctm.PEPSParameters(CHI=10, Max_CTM_Iteration=100, CTM_Convergence_Epsilon=1e-10,
                               CTM_Projector_corner=corner)
        env, count, _ = CS.Calc_CTM_Environment_density(Tn, p, lat)
        assert count < 100
        one, _ = itps.measure_onesite(env, Tn, lat, case["onesite"], K=KD)
        two, _ = itps.measure_twosite(env, Tn, lat, case["twosite"], case["onesite"], K=KD)
        vals[corner] = np.array([one[g][i] for g in sorted(one) for i in sorted(one[g])]
                                + [two[g][k] for g in sorted(two) for k in sorted(two[g])])
    assert np.max(np.abs(vals[True] - vals[False])) < tol
