"""Whole CTM runs on unit cells other than 2x2 (Calc_CTM_Environment,
src/iTPS/core/ctm.cpp:971-1253, with the lattice of src/SquareLattice.hpp):
1x1, 3x2, 2x3, 3x3 and skewed 3x2 / 2x1 cells, where the left/right and
top/bottom move chains pair different columns and rows than on the 2x2 cell,
with corner projectors and with the up-down projectors
(Calc_projector_updown_blocks, ctm.cpp:320-491; moves in sequence).  Seeded
random site tensors (D=2, unit norm) on the GPU and in the oracle, 4 sweeps
at chi=8: L2-normalised corner spectra of every site to 1e-9 and
nearest-neighbour Sz Sz bond values to 1e-9 relative."""
import numpy as np
import pytest

import tenes_amd as ta
from oracle import ctm as O, itps, models
from oracle.lattice import SquareLattice

from ctm_cases import random_tensors, spectra_diff

pytestmark = pytest.mark.gpu

CELLS = [(1, 1, 0), (3, 2, 0), (2, 3, 0), (3, 3, 0), (3, 2, 1), (2, 1, 1)]


@pytest.mark.parametrize("corner", [1, 0], ids=["corner", "updown"])
@pytest.mark.parametrize("LX,LY,skew", CELLS)
def test_ctm_on_unit_cell(gpu_ctx, LX, LY, skew, corner):
    lat = SquareLattice(LX, LY, skew)
    Tn = random_tensors(2, 2, seed=100 + 10 * LX + LY + skew, n=lat.N_UNIT)
    p = O.PEPSParameters(CHI=8, Max_CTM_Iteration=4, CTM_Convergence_Epsilon=0.0,
                         CTM_Projector_corner=bool(corner))
    env, count, _ = O.Calc_CTM_Environment(list(Tn), p, lat)
    genv = ta.Environment(gpu_ctx, LX, LY, skew)
    genv.set_site_tensors(Tn)
    gcount, _ = genv.calc_ctm_environment(
        ta.peps_parameters(CHI=8, Max_CTM_Iteration=4, CTM_Convergence_Epsilon=0.0,
                           CTM_Projector_corner=corner))
    assert gcount == count
    for i in range(lat.N_UNIT):
        for k in ("C1", "C2", "C3", "C4"):
            a = O.normalized_spectrum(getattr(env, k)[i])
            b = O.normalized_spectrum(genv.get(k, i).numpy())
            assert spectra_diff(a, b) < 1e-9, (k, i)
    SzSz = np.kron(models.Sz, models.Sz).reshape(2, 2, 2, 2)
    one = [{"name": "Sz", "group": 0, "site": i, "op": models.Sz} for i in range(lat.N_UNIT)]
    two = [{"name": "SzSz", "group": 0, "source": s, "dx": dx, "dy": dy, "op": SzSz}
           for s in range(lat.N_UNIT) for dx, dy in ((1, 0), (0, 1))]
    w, _ = itps.measure_twosite(env, list(Tn), lat, two, one)
    for t in two:
        s, dx, dy = t["source"], t["dx"], t["dy"]
        ref = float(np.real(w[0][(s, dx, dy)]))
        got = genv.contract_two_sites_op12(s, dx, dy, SzSz) / genv.contract_cluster(s, dx, dy)
        assert abs(got - ref) <= 1e-9 * max(abs(ref), 1e-3), (s, dx, dy, got, ref)
