"""GPU CTMRG path against the oracle on the same seeded inputs.

Renormalisation steps without an SVD (Calc_Next_CTM, Calc_Next_eT, the
environment contractions, the transfer matrix) are compared entry by entry.
Steps whose output depends on the SVD gauge (projectors, moves, whole
sweeps) are compared through gauge-invariant quantities: the oblique
projector PU*PL^T, L2-normalised corner spectra and observables.
Tolerances: 1e-11 relative for direct FP64 results, 1e-10 for observables
after whole CTM runs (FP64 mode target of the path)."""
import numpy as np
import pytest

import tenes_amd as ta
from oracle import contract as K, ctm as O, itps, transfer as TMo
from oracle.lattice import SquareLattice

from ctm_cases import random_tensors, spectra_diff, su_tensors
from oracle import models

pytestmark = pytest.mark.gpu

KINDS = ("C1", "C2", "C3", "C4", "eTt", "eTr", "eTb", "eTl")


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def random_env_pieces(rng, chi, D, d=2, cut=None):
    cut = cut or chi
    eT = rng.standard_normal((chi, chi, D, D))
    Tn = rng.standard_normal((D, D, D, D, d))
    PU = rng.standard_normal((chi, D, D, cut))
    PL = rng.standard_normal((chi, D, D, cut))
    return eT, Tn, PU, PL


@pytest.mark.parametrize("chi,D,cut", [(6, 2, 5), (20, 3, 18), (100, 5, 100)])
def test_calc_next_et(gpu_ctx, chi, D, cut):
    rng = np.random.default_rng(chi + D)
    eT, Tn, PU, PL = random_env_pieces(rng, chi, D, cut=cut)
    ref = O.Calc_Next_eT(eT, Tn, PU, PL)
    g = ta.calc_next_et(gpu_ctx.tensor(eT), gpu_ctx.tensor(Tn), gpu_ctx.tensor(PU), gpu_ctx.tensor(PL))
    assert g.shape == ref.shape
    assert rel(g.numpy(), ref) < 1e-11


@pytest.mark.parametrize("chi,D,cut", [(6, 2, 5), (40, 4, 33), (100, 5, 100), (30, 5, 27)])
def test_calc_next_et_fused_and_unfused(gpu_ctx, chi, D, cut):
    """Calc_Next_eT with its two site contractions fused into one kernel (the
    default) and as two gather-GEMMs: both against the oracle to 1e-11 and
    against each other to rounding.  The fused kernel takes the blocks of
    the cut leg four at a time: cut = 33 and 27 end on a group of 1 and 3."""
    rng = np.random.default_rng(5 * chi + D)
    eT, Tn, PU, PL = random_env_pieces(rng, chi, D, cut=cut)
    ref = O.Calc_Next_eT(eT, Tn, PU, PL)
    t = gpu_ctx.tensor
    out = []
    for fused in (True, False):
        gpu_ctx.set_fused_site(fused)
        try:
            out.append(ta.calc_next_et(t(eT), t(Tn), t(PU), t(PL)).numpy())
        finally:
            gpu_ctx.set_fused_site(True)
    for g in out:
        assert rel(g, ref) < 1e-11
    assert rel(out[0], out[1]) < 1e-13


@pytest.mark.parametrize("D,chi,sweeps,spec_tol", [(3, 20, 4, 1e-10), (5, 100, 2, 1e-9)])
def test_fused_site_ctm_matches_unfused(D, chi, sweeps, spec_tol):
    """Whole CTM runs (Heisenberg simple-update tensors) with the enlarged
    corners and the edge renormalisation fused and unfused, same seed: the
    arithmetic differs only in summation order, so energies agree to 1e-12;
    corner spectra to 1e-10, and to 1e-9 at D=5, chi=100, where the kept
    spectrum reaches 7.5e-7 of its largest value and the range finder's
    residual there (its s_129 / s_100 = 0.42 per iteration) carries the
    rounding difference into the smallest entries (measured 3.1e-10)."""
    Tn, ham = su_tensors("heisenberg", D, 60)
    lat = SquareLattice(2, 2)
    res = []
    for fused in (True, False):
        ctx = ta.Context(0, seed=17)
        ctx.set_fused_site(fused)
        genv = ta.Environment(ctx, 2, 2)
        genv.set_site_tensors(list(Tn))
        genv.calc_ctm_environment(ta.peps_parameters(CHI=chi, Max_CTM_Iteration=sweeps,
                                                     CTM_Convergence_Epsilon=0.0))
        E, _ = observables_gpu(genv, lat, ham)
        res.append((E, [genv.corner_spectrum(k, i) for i in range(4) for k in ("C1", "C2", "C3", "C4")]))
    (E1, s1), (E0, s0) = res
    assert abs(E1 - E0) < 1e-12 * abs(E0), (E1, E0)
    assert max(spectra_diff(a, b) for a, b in zip(s1, s0)) < spec_tol


@pytest.mark.parametrize("chi,D,cut", [(6, 2, 5), (40, 4, 33), (100, 5, 100)])
def test_calc_next_ctm(gpu_ctx, chi, D, cut):
    rng = np.random.default_rng(3 * chi + D)
    C1 = rng.standard_normal((chi, chi))
    C4 = rng.standard_normal((chi, chi))
    eT1 = rng.standard_normal((chi, chi, D, D))
    eT6 = rng.standard_normal((chi, chi, D, D))
    PU = rng.standard_normal((chi, D, D, cut))
    PL = rng.standard_normal((chi, D, D, cut))
    r1, r4 = O.Calc_Next_CTM(C1, C4, eT1, eT6, PU, PL)
    t = gpu_ctx.tensor
    g1, g4 = ta.calc_next_ctm(t(C1), t(C4), t(eT1), t(eT6), t(PU), t(PL))
    assert rel(g1.numpy(), r1) < 1e-11
    assert rel(g4.numpy(), r4) < 1e-11


def oracle_env(model, D, chi, steps, sweeps):
    Tn, ham = su_tensors(model, D, steps)
    Tn = list(Tn)
    lat = SquareLattice(2, 2)
    p = O.PEPSParameters(CHI=chi, Max_CTM_Iteration=sweeps, CTM_Convergence_Epsilon=0.0)
    env, _, _ = O.Calc_CTM_Environment(Tn, p, lat)
    return Tn, lat, env, ham


def upload_env(ctx, Tn, env):
    genv = ta.Environment(ctx, 2, 2)
    for i in range(4):
        genv.set("Tn", i, Tn[i])
        for k in KINDS:
            genv.set(k, i, getattr(env, k)[i])
    return genv


@pytest.mark.parametrize("model,D,chi,cut", [("heisenberg", 2, 10, 1e-12), ("tfi", 3, 20, 1e-10)])
def test_projector_left_block(gpu_ctx, model, D, chi, cut):
    """PU PL^T against the oracle.  The projector scales the kept singular
    vectors by s^-1/2, so any two FP64 SVDs (backward error ~eps*s0) give
    projectors that differ by up to ~eps*s0/s_min.  The tfi case has a
    singular value at s/s0 = 6.7e-12, just above the default 1e-12 cut,
    where that bound is 1.6e-5; it is compared with the cut at 1e-10 (smallest
    kept s/s0 = 3.7e-10).  Whole-run parity at the default cut:
    test_ctm_environment_observables."""
    Tn, lat, env, _ = oracle_env(model, D, chi, 40, 2)
    p = O.PEPSParameters(CHI=chi, Inverse_projector_cut=cut)
    i, l = 0, lat.bottom(0)
    args = (env.C1[i], env.C4[l], env.eTt[i], env.eTb[l], env.eTl[l], env.eTl[i], Tn[i], Tn[l])
    PU, PL = O.Calc_projector_left_block(*args, p)
    gp = ta.peps_parameters(CHI=chi, Inverse_projector_cut=cut)
    gPU, gPL = ta.calc_projector_left_block(*[gpu_ctx.tensor(a) for a in args], gp)
    assert gPU.shape == PU.shape and gPL.shape == PL.shape
    # the SVD gauge cancels in the oblique projector PU PL^T
    P = np.tensordot(PU, PL, ([3], [3]))
    gP = np.tensordot(gPU.numpy(), gPL.numpy(), ([3], [3]))
    assert rel(gP, P) < 1e-10


@pytest.mark.parametrize("D,chi", [(4, 7), (5, 5)])
def test_projector_left_block_fused_partial_groups(gpu_ctx, D, chi):
    """The enlarged corners through the fused site kernel on the MFMA (D = 4,
    5) when chi is not a multiple of its 4-block groups (chi = 7: a last group
    of 3; chi = 5: of 1), random environment pieces, chi D^2 <= 128 (exact
    Jacobi SVD): PU PL^T fused against unfused to 1e-12 and against the
    oracle to 1e-10."""
    rng = np.random.default_rng(11 * D + chi)
    C1, C4 = rng.standard_normal((chi, chi)), rng.standard_normal((chi, chi))
    eTt, eTb, eTl1, eTl0 = (rng.standard_normal((chi, chi, D, D)) for _ in range(4))
    Tn0, Tn1 = (rng.standard_normal((D, D, D, D, 2)) for _ in range(2))
    args = (C1, C4, eTt, eTb, eTl1, eTl0, Tn0, Tn1)
    PU, PL = O.Calc_projector_left_block(*args, O.PEPSParameters(CHI=chi))
    P = np.tensordot(PU, PL, ([3], [3]))
    gps = []
    for fused in (True, False):
        gpu_ctx.set_fused_site(fused)
        try:
            gPU, gPL = ta.calc_projector_left_block(*[gpu_ctx.tensor(a) for a in args],
                                                    ta.peps_parameters(CHI=chi))
        finally:
            gpu_ctx.set_fused_site(True)
        gps.append(np.tensordot(gPU.numpy(), gPL.numpy(), ([3], [3])))
    assert rel(gps[0], gps[1]) < 1e-12
    for gP in gps:
        assert rel(gP, P) < 1e-10


def test_projector_updown_blocks(gpu_ctx):
    """The kept spectrum of this case reaches s/s0 ~ 3e-12, where the singular
    vectors of any FP64 SVD carry ~eps*s0/s errors that the s^-1/2 scaling
    amplifies in PU PL^T; compare with a cutoff that keeps the well-conditioned
    part (s/s0 > 1e-6).  Whole-run parity of this projector: see
    test_ctm_environment_updown."""
    Tn, lat, env, _ = oracle_env("heisenberg", 2, 8, 40, 2)
    p = O.PEPSParameters(CHI=8, CTM_Projector_corner=False, Inverse_projector_cut=1e-6)
    i = 0
    j = lat.right(i)
    k = lat.bottom(j)
    l = lat.left(k)
    args = (env.C1[i], env.C2[j], env.C3[k], env.C4[l], env.eTt[i], env.eTt[j], env.eTr[j], env.eTr[k],
            env.eTb[k], env.eTb[l], env.eTl[l], env.eTl[i], Tn[i], Tn[j], Tn[k], Tn[l])
    PU, PL = O.Calc_projector_updown_blocks(*args, p)
    gPU, gPL = ta.calc_projector_updown_blocks([gpu_ctx.tensor(a) for a in args],
                                               ta.peps_parameters(CHI=8, CTM_Projector_corner=0,
                                                                  Inverse_projector_cut=1e-6))
    P = np.tensordot(PU, PL, ([3], [3]))
    gP = np.tensordot(gPU.numpy(), gPL.numpy(), ([3], [3]))
    assert rel(gP, P) < 1e-10


def observables_gpu(genv, lat, ham):
    E = 0.0
    for s, leg in models.square_bonds(lat):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        E += genv.contract_two_sites_op12(s, dx, dy, ham) / genv.contract_cluster(s, dx, dy)
    mz = [genv.contract_one_site(i, models.Sz) / genv.contract_one_site(i) for i in range(4)]
    return E / 4, np.array(mz)


def observables_oracle(env, Tn, lat, ham):
    one = [{"name": "Sz", "group": 0, "site": i, "op": models.Sz} for i in range(4)]
    two = []
    for s, leg in models.square_bonds(lat):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        two.append({"name": "bond_hamiltonian", "group": 0, "source": s, "dx": dx, "dy": dy, "op": ham})
    o, _ = itps.measure_onesite(env, Tn, lat, one)
    w, _ = itps.measure_twosite(env, Tn, lat, two, one)
    E = np.real(sum(w[0].values())) / 4
    return E, np.real(np.array([o[0][i] for i in range(4)]))


def test_environment_contractions_same_env(gpu_ctx):
    """One-site, two-site op12, 1x2/2x1 clusters and the transfer matrix on
    an identical (uploaded) environment: entry-level agreement."""
    Tn, lat, env, ham = oracle_env("heisenberg", 3, 12, 60, 4)
    genv = upload_env(gpu_ctx, Tn, env)
    E, mz = observables_oracle(env, Tn, lat, ham)
    gE, gmz = observables_gpu(genv, lat, ham)
    assert abs(gE - E) < 1e-12 * abs(E)
    assert np.max(np.abs(gmz - mz)) < 1e-12
    # product operators on a vertical bond through the 2x1 cluster
    v = genv.contract_cluster(0, 0, 1, models.Sz, models.Sz)
    nrow, ncol, sr, sc, tr, tc, idx = itps.cluster(lat, 0, 0, 1)
    C, eTt, eTr, eTb, eTl = itps._cluster_args(env, idx, nrow, ncol)
    ops = [[np.eye(2)] for _ in range(2)]
    ops[sr][sc] = models.Sz
    ops[tr][tc] = models.Sz
    ref = K.Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, [[Tn[idx[r][0]]] for r in range(2)], ops)
    assert abs(v - ref.real) < 1e-12 * abs(ref)
    # transfer matrix: dense (chi^2 = 144 <= 200) eigenvalues
    tm = TMo.TransferMatrixCTM(lat, Tn, env)
    for d in (0, 1):
        ref = tm.eigenvalues(d, 0, TMo.TMParameters())
        got = genv.transfer_matrix_eigenvalues(d, 0)
        assert rel(np.abs(got), np.abs(ref)) < 1e-11


def test_transfer_matrix_arnoldi(gpu_ctx):
    """chi^2 > 200: the Arnoldi path, compared with the oracle's Arnoldi and
    with a dense eigendecomposition of the whole transfer matrix."""
    Tn, lat, env, _ = oracle_env("heisenberg", 2, 20, 60, 6)
    genv = upload_env(gpu_ctx, Tn, env)
    tm = TMo.TransferMatrixCTM(lat, Tn, env)
    prm = TMo.TMParameters(num_eigvals=4)
    for d in (0, 1):
        ref = tm.eigenvalues(d, 0, prm)
        dense = TMo.eigen(tm.matrix(d, 0), 4)[0]
        got = genv.transfer_matrix_eigenvalues(d, 0, ta.tm_parameters(num_eigvals=4))
        assert rel(np.abs(got), np.abs(ref)) < 1e-9
        # leading eigenvalue is converged to the dense answer
        assert abs(abs(got[0]) - abs(dense[0])) < 1e-10 * abs(dense[0])


@pytest.mark.parametrize("maxdim,restartdim,maxiter", [(12, 8, 3), (12, 8, 10), (16, 6, 10),
                                                         (10, 4, 20)])
def test_transfer_matrix_arnoldi_restarts(gpu_ctx, maxdim, restartdim, maxiter):
    """Restarted Arnoldi (arnoldi.cpp:56-167, arnoldi_maxiter > 1) on the
    chi = 20 transfer matrix (N = 400) with a basis too small to converge in
    one pass: the 4 eigenvalues agree with the oracle's restarted run to 1e-8
    (moduli, relative; 1e-15 noise in the oracle's matvec moves them by at most
    5e-10 here) and, after 10 passes that keep more directions than
    eigenvalues, with the dense eigenvalues to 1e-11 (one pass of a 12-vector
    basis leaves them at 2e-4).  With restartdim = num_eigvals (10, 4, 20) the
    column before the first recomputed one misses the residual term the
    restart drops (as in the reference) and the run stalls at 4e-7 of the
    dense values, in the oracle alike."""
    Tn, lat, env, _ = oracle_env("heisenberg", 2, 20, 60, 6)
    genv = upload_env(gpu_ctx, Tn, env)
    tm = TMo.TransferMatrixCTM(lat, Tn, env)
    kw = dict(num_eigvals=4, arnoldi_maxdim=maxdim, arnoldi_restartdim=restartdim,
              arnoldi_maxiter=maxiter)
    for d in (0, 1):
        ref = tm.eigenvalues(d, 0, TMo.TMParameters(**kw))
        got = genv.transfer_matrix_eigenvalues(d, 0, ta.tm_parameters(**kw))
        assert np.max(np.abs(np.abs(got) - np.abs(ref)) / np.abs(ref)) < 1e-8, (got, ref)
        if maxiter >= 10 and restartdim > 4:
            dense = TMo.eigen(tm.matrix(d, 0), 4)[0]
            assert np.max(np.abs(np.abs(got) - np.abs(dense)) / np.abs(dense)) < 1e-11
    with pytest.raises(ta.TenesError, match="exceeds arnoldi_maxdim"):
        genv.transfer_matrix_eigenvalues(0, 0, ta.tm_parameters(num_eigvals=6, arnoldi_maxdim=5))


@pytest.mark.parametrize("model,D,chi,sweeps", [("heisenberg", 2, 10, 6), ("tfi", 2, 16, 5),
                                                ("heisenberg", 3, 20, 4)])
def test_ctm_environment_observables(gpu_ctx, model, D, chi, sweeps):
    """Calc_CTM_Environment from the same site tensors: energies, <Sz> and
    normalised corner spectra agree with the oracle."""
    Tn, lat, env, ham = oracle_env(model, D, chi, 60, sweeps)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=chi, Max_CTM_Iteration=sweeps,
                                                        CTM_Convergence_Epsilon=0.0))
    assert n == sweeps
    E, mz = observables_oracle(env, Tn, lat, ham)
    gE, gmz = observables_gpu(genv, lat, ham)
    assert abs(gE - E) < 1e-10 * abs(E)
    assert np.max(np.abs(gmz - mz)) < 1e-10
    for i in range(4):
        for k in ("C1", "C2", "C3", "C4"):
            ref = O.normalized_spectrum(getattr(env, k)[i])
            got = genv.corner_spectrum(k, i)
            assert spectra_diff(got, ref) < 1e-10


@pytest.mark.parametrize("device_cut", [1, 0])
def test_product_state_closed_form(gpu_ctx, device_cut):
    """Embedded product state (nonzero only at virtual index 0): the
    environment is exact at any D and chi, <S> and the bond energy take their
    product-state values (D=5, chi=100 shapes).  Every projector keeps one
    direction (s_1/s_0 = 0): with the cut on the host the environment shrinks
    to the reference's chi_eff = 1; with the cut on the device it stays
    100 x 100 with zeros outside that direction, and the observables agree."""
    v = np.array([np.cos(0.3), np.sin(0.3)])
    Tn = []
    for _ in range(4):
        T = np.zeros((5, 5, 5, 5, 2))
        T[0, 0, 0, 0] = v
        Tn.append(T)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=100, Max_CTM_Iteration=3, CTM_Convergence_Epsilon=0.0,
                                                 projector_cut_on_device=device_cut))
    ham = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    E, mz = observables_gpu(genv, SquareLattice(2, 2), ham)
    sz = v @ models.Sz @ v
    sx = v @ models.Sx @ v
    assert abs(mz[0] - sz) < 1e-13
    # S.S on |v v>: <Sz>^2 + <Sx>^2 (+ <Sy>^2 = 0 for a real v), 2 bonds/site
    assert abs(E - 2 * (sz * sz + sx * sx)) < 1e-13
    C = genv.get("C1", 0).numpy()
    assert C.shape == ((100, 100) if device_cut else (1, 1))
    s = np.linalg.svd(C, compute_uv=False)
    assert s[0] > 0 and np.all(s[1:] <= 1e-14 * s[0])


@pytest.mark.parametrize("model,D,chi,sweeps,floor", [("heisenberg", 2, 10, 6, 0.0), ("heisenberg", 3, 20, 4, 0.0),
                                                      ("tfi", 3, 20, 4, 1e-5)])
def test_device_and_host_cut_agree(gpu_ctx, model, D, chi, sweeps, floor):
    """The projector cut decided on the device (fixed CHI shapes, no host
    round trip) and on the host (the reference's shapes): energies, <Sz> and
    corner spectra agree to 1e-10, both with the oracle.  The TFI D=3, chi=20
    environment has projector singular values at s/s0 ~ 7e-12, next to the
    1e-12 cut, where any two FP64 SVDs may keep or drop the direction
    (test_projector_left_block): a dropped one carries ~3e-6 of a corner
    spectrum and nothing measurable of the observables, so there the spectra
    are compared above 1e-5."""
    Tn, lat, env, ham = oracle_env(model, D, chi, 60, sweeps)
    E, mz = observables_oracle(env, Tn, lat, ham)
    for dev in (1, 0):
        genv = ta.Environment(gpu_ctx, 2, 2)
        genv.set_site_tensors(Tn)
        n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=chi, Max_CTM_Iteration=sweeps,
                                                            CTM_Convergence_Epsilon=0.0,
                                                            projector_cut_on_device=dev))
        assert n == sweeps
        gE, gmz = observables_gpu(genv, lat, ham)
        assert abs(gE - E) < 1e-10 * abs(E), (dev, gE, E)
        assert np.max(np.abs(gmz - mz)) < 1e-10, dev
        for i in range(4):
            for k in ("C1", "C2", "C3", "C4"):
                a, b = genv.corner_spectrum(k, i), O.normalized_spectrum(getattr(env, k)[i])
                a, b = a[a > floor], b[b > floor]
                assert spectra_diff(a, b) < 1e-10, (dev, k, i)


def test_random_tensors_spectra(gpu_ctx):
    """Random unit-norm tensors (D=3, chi=27): a few full sweeps, corner
    spectra against the oracle."""
    Tn = random_tensors(3, seed=21)
    lat = SquareLattice(2, 2)
    chi = 27
    p = O.PEPSParameters(CHI=chi, Max_CTM_Iteration=3, CTM_Convergence_Epsilon=0.0)
    env, _, _ = O.Calc_CTM_Environment(Tn, p, lat)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=chi, Max_CTM_Iteration=3, CTM_Convergence_Epsilon=0.0))
    worst = 0.0
    for i in range(4):
        for k in ("C1", "C2", "C3", "C4"):
            ref = O.normalized_spectrum(getattr(env, k)[i])
            got = genv.corner_spectrum(k, i)
            worst = max(worst, np.max(np.abs(got - ref)))
    assert worst < 1e-10


def test_ctm_environment_updown(gpu_ctx):
    """Calc_CTM_Environment with the up-down (4-block) projector."""
    Tn, ham = su_tensors("heisenberg", 2, 60)
    Tn = list(Tn)
    lat = SquareLattice(2, 2)
    p = O.PEPSParameters(CHI=10, Max_CTM_Iteration=5, CTM_Convergence_Epsilon=0.0, CTM_Projector_corner=False)
    env, _, _ = O.Calc_CTM_Environment(Tn, p, lat)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=10, Max_CTM_Iteration=5, CTM_Convergence_Epsilon=0.0,
                                                 CTM_Projector_corner=0))
    E, mz = observables_oracle(env, Tn, lat, ham)
    gE, gmz = observables_gpu(genv, lat, ham)
    assert abs(gE - E) < 1e-10 * abs(E)
    assert np.max(np.abs(gmz - mz)) < 1e-10


@pytest.mark.parametrize("corner", [1, 0])
def test_rsvd_matrix_free_range_finder(gpu_ctx, corner):
    """The Use_RSVD branch at a size where the projector matrices are wider
    than 128 (unit-norm random tensors, D = 4, chi = 40: 640 x 640), so the
    range finder runs on the matrix-free operators of ctm.cpp:59-172 —
    LT (LB X) for the corner projectors, RT (LT (LB (RB X))) for the up-down
    ones, and the up-down projectors through the same block chains
    (ctm.cpp:400-439) — without forming R.  With 12 power iterations the
    randomized SVD equals the exact one to rounding, so the run equals the
    oracle's (exact SVD with the branch's cut): energy and <Sz> to 1e-10,
    corner spectra to 1e-9.  (Random tensors: a symmetric state's degenerate
    multiplets split by the cut at chi would leave the kept basis, and with
    it the corner spectra, to the SVD's choice.)"""
    Tn = random_tensors(4)
    ham = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    lat = SquareLattice(2, 2)
    p = O.PEPSParameters(CHI=40, Use_RSVD=True, CTM_Projector_corner=bool(corner), Max_CTM_Iteration=3,
                         CTM_Convergence_Epsilon=0.0)
    env, count, _ = O.Calc_CTM_Environment(list(Tn), p, lat)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(list(Tn))
    gcount, _ = genv.calc_ctm_environment(ta.peps_parameters(
        CHI=40, Use_RSVD=1, CTM_Projector_corner=corner, Max_CTM_Iteration=3, CTM_Convergence_Epsilon=0.0,
        rsvd_power_iters=12))
    assert gcount == count == 3
    E, mz = observables_oracle(env, Tn, lat, ham)
    gE, gmz = observables_gpu(genv, lat, ham)
    assert abs(gE - E) < 1e-10 * abs(E), (gE, E)
    assert np.max(np.abs(gmz - mz)) < 1e-10
    for i in range(4):
        for k in ("C1", "C2", "C3", "C4"):
            a = O.normalized_spectrum(getattr(env, k)[i])
            b = O.normalized_spectrum(genv.get(k, i).numpy())
            assert spectra_diff(a, b) < 1e-9, (k, i, spectra_diff(a, b))


def test_ctm_environment_rsvd_updown(gpu_ctx):
    """The Use_RSVD branch with up-down projectors (RSVD.toml's CTM settings:
    chi = 5, projector_corner = false; ctm.cpp:401-439), whose oracle is pinned
    to TeNeS's output_RSVD: on real simple-update tensors of that input the
    GPU run equals the oracle's to 1e-10 (the RSVD cut min(CHI, e34 t23^2),
    the 4-block projectors, energies and corner spectra)."""
    import golden_io as G
    from oracle import simple_update as SU

    fx = G.load("RSVD")
    LX, LY, skew, pdims, vdims, dirs, noises = G.unitcell(fx)
    lat = SquareLattice(LX, LY, skew)
    Tn = SU.initial_tensors(lat, pdims, vdims, dirs, noises, seed=11, is_real=True)
    lam = [[np.ones(vdims[i][l]) for l in range(4)] for i in range(lat.N_UNIT)]
    Tn, lam = SU.simple_update_steps(Tn, lam, lat, G.evolution(fx, True), 100)
    p = O.PEPSParameters(CHI=5, Use_RSVD=True, CTM_Projector_corner=False)
    env, count, _ = O.Calc_CTM_Environment(Tn, p, lat)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    gcount, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=5, Use_RSVD=1, CTM_Projector_corner=0))
    assert gcount == count
    for i in range(4):
        for k in ("C1", "C2", "C3", "C4"):
            a = O.normalized_spectrum(getattr(env, k)[i])
            b = O.normalized_spectrum(genv.get(k, i).numpy())
            assert spectra_diff(a, b) < 1e-10, (k, i)
    one, _, two, _ = G.operators(fx, 4, True)
    o, _ = itps.measure_onesite(env, Tn, lat, one)
    w, _ = itps.measure_twosite(env, Tn, lat, two, one)
    for x in one:
        got = genv.contract_one_site(x["site"], x["op"]) / genv.contract_one_site(x["site"])
        assert abs(got - o[x["group"]][x["site"]]) < 1e-10, x["name"]
    for x in two:
        if "op" not in x:
            continue
        got = genv.contract_two_sites_op12(x["source"], x["dx"], x["dy"], x["op"]) / \
            genv.contract_cluster(x["source"], x["dx"], x["dy"])
        assert abs(got - w[x["group"]][(x["source"], x["dx"], x["dy"])]) < 1e-10


def test_simple_update_bond_gpu(gpu_ctx):
    """Simple_update_bond: bond weights (gauge invariant) and the two-site
    wavefunction contracted over the updated bond against the oracle."""
    from oracle import simple_update as SU

    rng = np.random.default_rng(4)
    D, d = 3, 2
    T1 = rng.standard_normal((D, D, D, D, d))
    T2 = rng.standard_normal((D, D, D, D, d))
    lam1 = [rng.uniform(0.2, 1.0, D) for _ in range(4)]
    lam2 = [rng.uniform(0.2, 1.0, D) for _ in range(4)]
    gate = models.evolution_twosite(models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0), 0.05)
    for leg in range(4):
        r1, r2, rl = SU.Simple_update_bond(T1, T2, lam1, lam2, gate, leg)
        g1, g2, gl = ta.simple_update_bond(gpu_ctx.tensor(T1), gpu_ctx.tensor(T2), lam1, lam2,
                                           gpu_ctx.tensor(gate), leg)
        assert np.max(np.abs(gl - rl)) < 1e-13
        l2 = (leg + 2) % 4
        ref = np.tensordot(r1, r2, ([leg], [l2]))
        got = np.tensordot(g1.numpy(), g2.numpy(), ([leg], [l2]))
        assert rel(got, ref) < 1e-12


def test_simple_update_steps_energy(gpu_ctx):
    """20 simple-update steps on the GPU from the same initial tensors as the
    oracle, then CTM on both: energies agree (bond gauges may differ)."""
    from oracle import simple_update as SU

    lat = SquareLattice(2, 2)
    D = 2
    Tn0 = SU.initial_tensors(lat, [2] * 4, [[D] * 4] * 4, models.neel_dirs(lat), [0.01] * 4, seed=11)
    ham = models.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    gate = models.evolution_twosite(ham, 0.01)
    ops = [(s, leg, gate) for s, leg in models.square_bonds(lat)]
    lam = [[np.ones(D) for _ in range(4)] for _ in range(4)]
    Tn, lam = SU.simple_update_steps([t.copy() for t in Tn0], lam, lat, ops, 20)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn0)
    genv.simple_update(ops, 20)
    for s in range(4):
        for leg in range(4):
            assert np.max(np.abs(genv.get_lambda(s, leg) - lam[s][leg])) < 1e-12
    p = O.PEPSParameters(CHI=8, Max_CTM_Iteration=4, CTM_Convergence_Epsilon=0.0)
    env, _, _ = O.Calc_CTM_Environment(Tn, p, lat)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=8, Max_CTM_Iteration=4, CTM_Convergence_Epsilon=0.0))
    E, _ = observables_oracle(env, Tn, lat, ham)
    gE, _ = observables_gpu(genv, lat, ham)
    assert abs(gE - E) < 1e-10 * abs(E)


def test_afh_real_goldens_on_gpu(gpu_ctx):
    """GPU CTM + GPU observables on the AntiferroHeisenberg_real site tensors
    (SU 100 + FU 1 prepared by the oracle) against TeNeS's own output:
    energy, <Sz>, <Sx>, SzSz/SxSx/SySy bond averages and correlation lengths."""
    import golden_io as G

    fx, lat, pdims, Tn = G.afh_real_site_tensors()
    chi = fx["parameter"]["ctm"]["dimension"]
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=chi))
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
    for name, v in vals.items():
        assert abs(v - gold[name].real) < 1e-10, (name, v, gold[name])
    assert abs(vals["bond_hamiltonian"] - gold["Energy"].real) < 1e-10
    for r in fx["outputs"]["correlation_length"]:
        ref = [float(x) for x in r.split()]
        d, f = int(ref[0]), int(ref[1])
        ev = genv.transfer_matrix_eigenvalues(d, f)
        L = 2
        xi = -L / np.log(abs(ev[1]) / abs(ev[0]))
        assert abs(xi - ref[2]) < 1e-10


def _fixture_or_skip(name):
    import golden_io as G
    try:
        return G.load(name)
    except FileNotFoundError:
        pytest.skip(f"fixture {name}.json not generated (tests/golden/make_oracle_runs.py)")


@pytest.mark.parametrize("kw", [{}, {"svd_power_iters": 16, "svd_warm_start": 0},
                                {"svd_power_iters": 16, "svd_warm_start": 1}],
                         ids=["library-defaults", "cold-q16", "warm-after-q16"])
def test_ctm_random_d4_chi64_30_sweeps(gpu_ctx, kw):
    """CTM-only case: 2x2, D=4, chi=64, unit-norm random site tensors, 30
    sweeps with epsilon = 0; L2-normalised spectra of every C1..C4 against the
    oracle's LAPACK-SVD run (tests/golden/ctm_random_d4_chi64.json) to 1e-10,
    at the library's default range-finder settings and with 16 iterations
    from a Gaussian start (cold throughout, or warm-started after the first
    sweep).  The spectrum of R decays slowly for random tensors (s_128/s_64 ~
    0.38, so the range finder's subspace error after q cold iterations is
    ~0.38^(2q+1)): without the warm start, q = 4 / 8 / 16 give 2.5e-7 /
    3.2e-10 / 4.4e-15.  The warm start accumulates iterations over the sweeps
    (the environment settles: corner-spectrum change 7e-8 at sweep 7)."""
    from tenes_amd import inputs

    fx = _fixture_or_skip("ctm_random_d4_chi64")
    Tn = inputs.random_tensors(4, fx["D"], 2, seed=fx["seed"])
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=fx["chi"], Max_CTM_Iteration=fx["sweeps"],
                                                        CTM_Convergence_Epsilon=0.0, **kw))
    assert n == fx["sweeps"]
    worst = 0.0
    for i in range(4):
        for k in ("C1", "C2", "C3", "C4"):
            ref = np.array(fx["spectra"][f"{k}{i}"])
            got = genv.corner_spectrum(k, i)
            assert got.shape == ref.shape
            worst = max(worst, float(np.max(np.abs(got - ref))))
    print(f"D=4 chi=64 30 sweeps {kw or 'library defaults'}: worst spectrum difference {worst:.2e}")
    assert worst < 1e-10, worst


def test_ctm_warm_start_rank_growth(gpu_ctx):
    """Warm-started range finder across a rank increase: D=3, chi=32 random
    tensors (projector matrices 288 x 288, so the randomized path).  The
    environment starts with chi_eff = D^2 = 9, so the first sweep's matrices
    have rank <= 81 < 128 and the Ritz basis carried to the next sweep has
    numerically null directions; they must be refilled or the range finder
    never sees the directions the grown environment adds.  Spectra after 10
    sweeps against the oracle to 1e-10 (16 iterations cold and warm), and the
    same run continued over two calls (initialize = false keeps the bases)."""
    Tn = random_tensors(3, 2, seed=17)
    lat = SquareLattice(2, 2)
    env, _, _ = O.Calc_CTM_Environment(list(Tn), O.PEPSParameters(CHI=32, Max_CTM_Iteration=10,
                                                                   CTM_Convergence_Epsilon=0.0), lat)
    kw = dict(CHI=32, CTM_Convergence_Epsilon=0.0, svd_power_iters=16, svd_warm_power_iters=16)
    for split in ((10,), (4, 6)):
        genv = ta.Environment(gpu_ctx, 2, 2)
        genv.set_site_tensors(Tn)
        for k, n in enumerate(split):
            genv.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=n, **kw), initialize=(k == 0))
        worst = 0.0
        for i in range(4):
            for kind in ("C1", "C2", "C3", "C4"):
                ref = O.normalized_spectrum(getattr(env, kind)[i])
                got = genv.corner_spectrum(kind, i)
                worst = max(worst, spectra_diff(got, ref))
        assert worst < 1e-10, (split, worst)


def test_afh_real_correlation_functions_on_gpu(gpu_ctx):
    """Correlation functions <Sz(0) Sz(r)>, <Sz(0) Sx(r)>, r = 1..3 along +x
    and +y from every site, and the four-site SzSzSzSz observables on the 2x2
    cluster, on the GPU CTM environment of the AntiferroHeisenberg_real site
    tensors, against TeNeS's correlation.dat (48 values) and
    multisite_obs_4.dat (8 values) to 1e-10."""
    import golden_io as G

    fx, lat, pdims, Tn = G.afh_real_site_tensors()
    chi = fx["parameter"]["ctm"]["dimension"]
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=chi))
    one, _, _, _ = G.operators(fx, lat.N_UNIT, True)
    site_ops = {(o["site"], o["group"]): o["op"] for o in one}
    r_max = fx["correlation"]["r_max"]
    pairs = [tuple(x) for x in fx["correlation"]["operators"]]
    gold = {}
    for r in fx["outputs"]["correlation"]:
        f = r.split()
        gold[(int(f[0]), int(f[1]), int(f[2]), int(f[3]), int(f[4]))] = float(f[5])
    n = 0
    for left in range(4):
        for lg in sorted({p[0] for p in pairs}):
            rgs = [p[1] for p in pairs if p[0] == lg]
            for direction in (0, 1):
                vals = genv.correlation(left, site_ops[(left, lg)], [site_ops[(left, rg)] for rg in rgs],
                                        r_max, direction)
                for r in range(r_max):
                    for j, rg in enumerate(rgs):
                        key = (lg, left, rg, r + 1 if direction == 0 else 0, r + 1 if direction == 1 else 0)
                        assert abs(vals[r, j] - gold[key]) < 1e-10, (key, vals[r, j], gold[key])
                        n += 1
    assert n == len(gold) == 48
    # four-site observables on the 2x2 cluster (multisite_obs_4.dat)
    rows = fx["outputs"]["multisite_obs_4"]
    n = 0
    for m in fx["multisite"]:
        for s in m["sites"]:
            src = s["source"]
            ops = [site_ops[(src, m["ops"][0])]] + [site_ops[(lat.other(src, x, y), g)]
                                                   for x, y, g in zip(s["dx"], s["dy"], m["ops"][1:])]
            v = m["coeff"] * genv.contract_multisite(src, s["dx"], s["dy"], ops)
            key = " ".join(str(x) for x in [m["group"], src] + [v_ for pair in zip(s["dx"], s["dy"]) for v_ in pair])
            ref = [float(r.split()[8]) for r in rows if r.startswith(key + " ")]
            assert len(ref) == 1 and abs(v - ref[0]) < 1e-10, (key, v, ref)
            n += 1
    assert n == len(rows) == 8


def test_heisenberg_d5_chi100_energy(gpu_ctx):
    """The benchmark problem (Heisenberg 2x2, D=5, chi=100) from the oracle's
    own simple-update site tensors: the GPU's fixed-sweep CTM energy per site,
    <Sz>, <Sx> of every site, the staggered magnetisation and the 10 leading
    transfer-matrix eigenvalues / correlation lengths (chi^2 = 10^4, Arnoldi)
    against the oracle's LAPACK-SVD run (tests/golden/heisenberg_d5_chi100.json),
    default SVD settings."""
    from tenes_amd import inputs

    fx = _fixture_or_skip("heisenberg_d5_chi100")
    D = fx["D"]
    Tn = [np.array(t).reshape(D, D, D, D, 2) for t in fx["site_tensors"]]
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=fx["chi"], Max_CTM_Iteration=fx["sweeps"],
                                                        CTM_Convergence_Epsilon=0.0))
    assert n == fx["sweeps"]
    H = inputs.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    E = 0.0
    for s, leg in inputs.square_bonds(2, 2):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        E += genv.contract_two_sites_op12(s, dx, dy, H) / genv.contract_cluster(s, dx, dy)
    E /= 4.0
    assert abs(E - fx["energy_per_site"]) < 1e-10 * abs(fx["energy_per_site"]), (E, fx["energy_per_site"])
    sz, sx = [], []
    for i in range(4):
        nrm = genv.contract_one_site(i)
        sz.append(genv.contract_one_site(i, models.Sz) / nrm)
        sx.append(genv.contract_one_site(i, models.Sx) / nrm)
        assert abs(sz[-1] - fx["sz"][i]) < 1e-10
        assert abs(sx[-1] - fx["sx"][i]) < 1e-10
    par = [1, -1, -1, 1]
    ms = np.hypot(np.dot(par, sz), np.dot(par, sx)) / 4.0
    assert abs(ms - fx["staggered_magnetization"]) < 1e-10 * fx["staggered_magnetization"]
    for c in fx["correlation"]:
        ev = genv.transfer_matrix_eigenvalues(c["dir"], c["fixed"], ta.tm_parameters(num_eigvals=10))
        ref = np.array(c["eigvals_re"]) + 1j * np.array(c["eigvals_im"])
        xi, _ = TMo.correlation_length(ev, 2)
        # leading eigenvalue: converged, 1e-10.  xi = -L/ln|t1/t0| follows
        # the environment's tail: with the range finder converged (16-24
        # iterations) the GPU's xi differ from the oracle's LAPACK run by
        # 2.6e-9, 2.3e-8, 9e-9, 1.2e-8 (the four entries); the default warm
        # start reaches 2.7e-9, 2.3e-8, 1.3e-8, 1.7e-8: 5e-8.  The deeper Ritz
        # values are not converged and follow rounding differences of the
        # environment: 1e-6
        assert abs(abs(ev[0]) - abs(ref[0])) < 1e-10 * abs(ref[0])
        assert abs(xi - c["xi"]) < 5e-8 * c["xi"], (xi, c["xi"])
        assert np.max(np.abs(np.abs(ev) - np.abs(ref)) / np.abs(ref[0])) < 1e-6


@pytest.mark.parametrize("mode,e_tol,xi_tol", [(1, 1e-7, 5e-5), (2, 5e-7, 5e-5)])
def test_heisenberg_d5_chi100_mixed_precision(mode, e_tol, xi_tol):
    """The benchmark problem with the contractions of the moves on the FP32
    MFMA (projector SVDs, QRs, convergence check, transfer-matrix Arnoldi and
    observables in FP64), against the oracle's FP64 LAPACK run.  Mode 1 keeps
    the projector matrix R = LT LB in FP64: energy within the mixed-mode
    target 1e-7 (measured 7.6e-9 to 2.8e-8 over range-finder seeds).  Mode 2
    puts R on the FP32 MFMA too: its smallest kept singular values (7.5e-7 of
    the largest) sit near the FP32 accumulation error of R, and the energy
    moves by 2.3e-7 to 2.5e-7 (bar 5e-7).  The correlation lengths (ratio of
    the two leading transfer-matrix eigenvalues, a gap of ~0.54) move by
    6e-7 to 1.4e-5 depending on the range finder's random start (bar 5e-5).
    A context of its own with a fixed seed, so the random starts do not
    depend on which tests ran before."""
    from tenes_amd import inputs

    fx = _fixture_or_skip("heisenberg_d5_chi100")
    D = fx["D"]
    Tn = [np.array(t).reshape(D, D, D, D, 2) for t in fx["site_tensors"]]
    ctx = ta.Context(0, seed=4242)
    genv = ta.Environment(ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=fx["chi"], Max_CTM_Iteration=fx["sweeps"],
                                                        CTM_Convergence_Epsilon=0.0, contraction_fp32=mode))
    assert n == fx["sweeps"]
    H = inputs.bond_hamiltonian(Jz=1.0, Jx=1.0, Jy=1.0)
    E = 0.0
    for s, leg in inputs.square_bonds(2, 2):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        E += genv.contract_two_sites_op12(s, dx, dy, H) / genv.contract_cluster(s, dx, dy)
    E /= 4.0
    rel = abs(E - fx["energy_per_site"]) / abs(fx["energy_per_site"])
    xis = []
    for c in fx["correlation"]:
        ev = genv.transfer_matrix_eigenvalues(c["dir"], c["fixed"], ta.tm_parameters(num_eigvals=10))
        xi, _ = TMo.correlation_length(ev, 2)
        xis.append(abs(xi - c["xi"]) / c["xi"])
    print(f"mixed precision {mode}: energy rel err {rel:.2e}, xi rel errs {['%.1e' % x for x in xis]}")
    assert rel < e_tol, rel
    assert max(xis) < xi_tol, xis


def test_tfi_plain_reference_config(gpu_ctx):
    """Transverse-field Ising (Jz = -1, hx = 1), 2x2, D = 2, chi = 10: 1000
    simple-update steps (tau = 0.01) from the ferro state on the GPU and
    TeNeS's CTM driver (10 iterations, default convergence epsilon) -- energy
    per site, <Sz> and <Sx> against the oracle running the same steps on the
    same seeded inputs."""
    from oracle import simple_update as SU

    lat = SquareLattice(2, 2)
    D = 2
    Tn0 = SU.initial_tensors(lat, [2] * 4, [[D] * 4] * 4, [[1.0, 0.0]] * 4, [0.01] * 4, seed=11)
    ham = models.bond_hamiltonian(Jz=-1.0, hx=1.0)
    gate = models.evolution_twosite(ham, 0.01)
    ops = [(s, leg, gate) for s, leg in models.square_bonds(lat)]
    lam = [[np.ones(D) for _ in range(4)] for _ in range(4)]
    Tn, lam = SU.simple_update_steps([t.copy() for t in Tn0], lam, lat, ops, 1000)
    p = O.PEPSParameters(CHI=10, Max_CTM_Iteration=10)
    env, count, _ = O.Calc_CTM_Environment(Tn, p, lat)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn0)
    genv.simple_update(ops, 1000)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=10, Max_CTM_Iteration=10))
    assert n == count
    E, mz = observables_oracle(env, Tn, lat, ham)
    gE, gmz = observables_gpu(genv, lat, ham)
    assert abs(gE - E) < 1e-10 * abs(E), (gE, E)
    assert np.max(np.abs(gmz - mz)) < 1e-10
    one = [{"name": "Sx", "group": 0, "site": i, "op": models.Sx} for i in range(4)]
    o, _ = itps.measure_onesite(env, Tn, lat, one)
    gmx = [genv.contract_one_site(i, models.Sx) / genv.contract_one_site(i) for i in range(4)]
    assert np.max(np.abs(np.array(gmx) - np.real([o[0][i] for i in range(4)]))) < 1e-10


def test_paired_move_chains_equal_reference_order(gpu_ctx):
    """Calc_CTM_Environment runs the left and right (then top and bottom)
    move chains side by side; the same sweeps issued one move at a time in
    the reference's order (ctm.cpp:1203-1228: left 0, 1; right 1, 0; top 1,
    0; bottom 0, 1) give the same environment: corner spectra and energy to
    1e-12 (D=2, chi=16: every projector matrix takes the exact Jacobi SVD)."""
    Tn, ham = su_tensors("heisenberg", 2, 60)
    Tn = list(Tn)
    lat = SquareLattice(2, 2)
    kw = dict(CHI=16, CTM_Convergence_Epsilon=0.0)
    a = ta.Environment(gpu_ctx, 2, 2)
    a.set_site_tensors(Tn)
    a.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=3, **kw))
    b = ta.Environment(gpu_ctx, 2, 2)
    b.set_site_tensors(Tn)
    n, _ = b.calc_ctm_environment(ta.peps_parameters(Max_CTM_Iteration=0, **kw))
    assert n == 0
    p = ta.peps_parameters(**kw)
    for _ in range(3):
        for d, coords in ((0, (0, 1)), (1, (1, 0)), (2, (1, 0)), (3, (0, 1))):
            for c in coords:
                b.move(d, c, p)
    for i in range(4):
        for k in ("C1", "C2", "C3", "C4"):
            sa = O.normalized_spectrum(a.get(k, i).numpy())
            sb = O.normalized_spectrum(b.get(k, i).numpy())
            assert sa.shape == sb.shape and np.max(np.abs(sa - sb)) < 1e-12, (k, i)
    Ea, _ = observables_gpu(a, lat, ham)
    Eb, _ = observables_gpu(b, lat, ham)
    assert abs(Ea - Eb) < 1e-12 * abs(Ea)
