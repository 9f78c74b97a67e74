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
