"""Cluster contractions of every shape up to 4x4 on the GPU
(Contract_iTPS_CTM, contract_itps_ctm/ctm.cpp:45-87 and ctm_NxM.cpp) against
the oracle's, through the C ABI: multisite expectation values
(tenes_env_contract_multisite, normalised by the identity cluster) and
two-site product clusters (tenes_env_contract_cluster, unnormalised).
Random FP64 environments; tolerance 1e-11 relative (1e-10 for ratios)."""
import numpy as np
import pytest

import tenes_amd as ta
from oracle import contract as K
from oracle.itps import _cluster_args
from oracle.lattice import SquareLattice

from test_oracle_cluster import random_env

pytestmark = pytest.mark.gpu

KINDS = ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl")
SHAPES = [(r, c) for r in range(1, 5) for c in range(1, 5)]


@pytest.fixture(scope="module")
def case(gpu_ctx):
    lat = SquareLattice(2, 2)
    env, Tn, ops = random_env(lat, 6, 2, 2, seed=21)
    genv = ta.Environment(gpu_ctx, 2, 2)
    for i in range(lat.N_UNIT):
        genv.set("Tn", i, Tn[i])
        for k in KINDS:
            genv.set(k, i, getattr(env, k)[i])
    rng = np.random.default_rng(5)
    site_ops = [rng.standard_normal((2, 2)) for _ in range(4)]
    return lat, env, Tn, genv, site_ops


def oracle_cluster(env, Tn, lat, src, offsets, ops):
    """<prod ops at src + offsets> over the smallest cluster holding them
    (the cluster rule of multisite_obs.cpp:52-155) and its identity norm."""
    dxs = [0] + [o[0] for o in offsets]
    dys = [0] + [o[1] for o in offsets]
    ncol, nrow = max(dxs) - min(dxs) + 1, max(dys) - min(dys) + 1
    scol, srow = -min(dxs), max(dys)
    idx = [[lat.other(src, c - scol, srow - r) for c in range(ncol)] for r in range(nrow)]
    C, eTt, eTr, eTb, eTl = _cluster_args(env, idx, nrow, ncol)
    T = [[Tn[idx[r][c]] for c in range(ncol)] for r in range(nrow)]
    O = [[np.eye(2) for _ in range(ncol)] for _ in range(nrow)]
    norm = K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, T, O).real
    for (dx, dy), op in zip([(0, 0)] + list(offsets), ops):
        O[srow - dy][scol + dx] = op
    return K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, T, O).real, norm


@pytest.mark.parametrize("shape", SHAPES)
def test_multisite_every_shape(case, shape):
    lat, env, Tn, genv, site_ops = case
    nrow, ncol = shape
    # operators on the corners of the cluster (source at the bottom left)
    offsets = sorted({(ncol - 1, 0), (0, nrow - 1), (ncol - 1, nrow - 1)} - {(0, 0)})
    if not offsets:
        offsets = []
    ops = site_ops[: 1 + len(offsets)]
    for src in range(lat.N_UNIT):
        val, norm = oracle_cluster(env, Tn, lat, src, offsets, ops)
        got = genv.contract_multisite(src, [o[0] for o in offsets], [o[1] for o in offsets], ops)
        assert abs(got - val / norm) <= 1e-10 * abs(val / norm), (shape, src, got, val / norm)


@pytest.mark.parametrize("dx,dy", [(2, 0), (-3, 0), (0, 3), (0, -2), (3, 3), (-2, 1), (1, -3)])
def test_two_site_product_clusters(case, dx, dy):
    lat, env, Tn, genv, site_ops = case
    A, B = site_ops[0], site_ops[1]
    for src in range(lat.N_UNIT):
        val, _ = oracle_cluster(env, Tn, lat, src, [(dx, dy)], [A, B])
        got = genv.contract_cluster(src, dx, dy, A, B)
        assert abs(got - val) <= 1e-11 * abs(val), (dx, dy, src, got, val)


def test_cluster_beyond_4x4_is_an_error(case):
    lat, env, Tn, genv, site_ops = case
    with pytest.raises(ta.TenesError, match="not implemented"):
        genv.contract_cluster(0, 4, 0, site_ops[0], site_ops[1])


def test_bad_source_site_is_an_error(case):
    lat, env, Tn, genv, site_ops = case
    with pytest.raises(ta.TenesError, match="bad source site"):
        genv.contract_cluster(4, 1, 0)
    with pytest.raises(ta.TenesError, match="bad source site"):
        genv.contract_two_sites_op12(-1, 1, 0, np.eye(4).reshape(2, 2, 2, 2))
    with pytest.raises(ta.TenesError, match="bad multisite operator"):
        genv.contract_multisite(7, [1], [0], [np.eye(2), np.eye(2)])
