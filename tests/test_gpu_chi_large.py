"""Bond dimensions above 112: the truncated SVD with a 256-column range
finder (Gram matrices factorised as two 128 blocks, one-sided Jacobi of the
256 x 256 projected problem in global memory) and corner spectra of
matrices above 128 x 128 (the convergence check).  The SVD is checked
against LAPACK (numpy), whole CTM runs against the oracle (numpy/LAPACK,
exact SVD), through L2-normalised corner spectra to 1e-10."""
import functools

import numpy as np
import pytest

import tenes_amd as ta
from oracle import ctm as O
from oracle.lattice import SquareLattice

from ctm_cases import random_tensors, spectra_diff

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def _oracle_run(seed, chi, sweeps, eps):
    Tn = random_tensors(2, 2, seed=seed)
    env, n, sig = O.Calc_CTM_Environment(list(Tn), O.PEPSParameters(CHI=chi, Max_CTM_Iteration=sweeps,
                                                                     CTM_Convergence_Epsilon=eps),
                                         SquareLattice(2, 2))
    return Tn, env, n, sig


def _check_svd(A, U, s, V, k, srel, orth=1e-10):
    s_ref = np.linalg.svd(A, compute_uv=False)[:k]
    assert s.shape == (k,)
    err = float(np.max(np.abs(s - s_ref)) / s_ref[0])
    assert err < srel, err
    assert np.max(np.abs(U.T @ U - np.eye(k))) < orth
    assert np.max(np.abs(V.T @ V - np.eye(k))) < orth
    assert np.max(np.abs(A @ V - U * s)) / s[0] < 1e-10


def _matrix(rng, n, sv):
    Q1, _ = np.linalg.qr(rng.standard_normal((n, sv.size)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, sv.size)))
    return (Q1 * sv) @ Q2.T


def test_svd_randomized_k160(gpu_ctx):
    """k = 160 of a 1500 x 1500 matrix with singular values exp(-i/25),
    i < 600: the 256-wide block (96 columns of oversampling, subspace error
    ~(s_257/s_160)^9 ~ 1e-15 after 4 iterations)."""
    rng = np.random.default_rng(21)
    A = _matrix(rng, 1500, np.exp(-np.arange(600) / 25.0))
    U, s, V = ta.svd_top(gpu_ctx.tensor(A), 160, power_iters=4)
    _check_svd(A, U.numpy(), s, V.numpy(), 160, 1e-12)


def test_svd_randomized_k240_gap(gpu_ctx):
    """k = 240, the largest with 16 columns of oversampling in the 256-wide
    block: 240 singular values from 1 down to 1e-6, the rest below 1e-9 (a
    gap of 1e3 bounds the subspace error)."""
    rng = np.random.default_rng(22)
    sv = np.concatenate([np.logspace(0, -6, 240), 1e-9 * np.exp(-np.arange(360) / 50.0)])
    A = _matrix(rng, 1200, sv)
    U, s, V = ta.svd_top(gpu_ctx.tensor(A), 240, power_iters=4)
    _check_svd(A, U.numpy(), s, V.numpy(), 240, 1e-12)


def test_svd_k_above_240_rejected(gpu_ctx):
    """k = 241 (fewer than 16 columns of oversampling in the 256-wide block)
    and k = 300 are refused with TENES_ERR_UNSUPPORTED."""
    A = np.random.default_rng(23).standard_normal((600, 600))
    with pytest.raises(ta.TenesError, match="error -5: .*range-finder block"):
        ta.svd_top(gpu_ctx.tensor(A), 241)
    with pytest.raises(ta.TenesError, match="error -5: .*at most 256"):
        ta.svd_top(gpu_ctx.tensor(A), 300)


def test_svd_randomized_k160_rank_deficient(gpu_ctx):
    """Rank 130 < 256: the second 128-column block of the range finder is
    numerically dependent on the first; the blocked factorisation drops those
    columns (relative pivot <= 1e-13 against the whole matrix, as the
    one-block factorisation does) and the singular values beyond the rank
    come out below 1e-12 of the largest."""
    rng = np.random.default_rng(24)
    n = 900
    A = rng.standard_normal((n, 130)) @ rng.standard_normal((130, n))
    U, s, V = ta.svd_top(gpu_ctx.tensor(A), 160)
    s_ref = np.linalg.svd(A, compute_uv=False)[:160]
    assert np.max(np.abs(s[:130] - s_ref[:130])) / s_ref[0] < 1e-12
    assert np.all(s[130:] < 1e-12 * s[0])
    Uk, Vk = U.numpy()[:, :130], V.numpy()[:, :130]
    assert np.max(np.abs(A @ Vk - Uk * s[:130])) / s[0] < 1e-10


@pytest.mark.parametrize("kw", [{}, {"svd_power_iters": 16, "svd_warm_start": 0}],
                         ids=["library-defaults", "cold-q16"])
def test_ctm_random_d2_chi160(gpu_ctx, kw):
    """2x2, D = 2, chi = 160, unit-norm random site tensors, 8 sweeps with
    epsilon = 0 (the corners reach 160 x 160 in the third sweep; projector
    matrices 640 x 640): corner spectra of every C1..C4 against the oracle's
    run, at the library's default range-finder settings and with 16 cold
    iterations.  Tolerance 1e-8, not 1e-10: the kept singular values of R go
    down to s_160 / s_1 ~ 4e-8 and the projectors scale by s^(-1/2), so two
    FP64 SVDs of R agree in those directions only to ~eps s_1 / s_k.
    Measured: 2.5e-9 with 16, 24 cold iterations alike and at the defaults
    (the floor; a numpy restatement of the range finder lands at 2.0e-9)."""
    sweeps = 8
    Tn, env, _, _ = _oracle_run(41, 160, sweeps, 0.0)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, _ = genv.calc_ctm_environment(ta.peps_parameters(CHI=160, Max_CTM_Iteration=sweeps,
                                                        CTM_Convergence_Epsilon=0.0, **kw))
    assert n == sweeps
    worst = 0.0
    for i in range(4):
        for kind in ("C1", "C2", "C3", "C4"):
            ref = O.normalized_spectrum(getattr(env, kind)[i])
            got = genv.corner_spectrum(kind, i)
            assert got.size == 160
            worst = max(worst, spectra_diff(got, ref))
    print(f"D=2 chi=160 {sweeps} sweeps {kw or 'library defaults'}: worst spectrum difference {worst:.2e}")
    assert worst < 1e-8, worst


def test_ctm_chi160_convergence_check(gpu_ctx):
    """The convergence test on 160 x 160 corners (batched Jacobi of 16
    matrices in global memory): with epsilon = 1e-30 no run converges, and
    the reported last change agrees with the oracle's convergence quantity
    after the same sweeps to 1e-8 relative."""
    sweeps = 6
    pp = dict(CHI=160, Max_CTM_Iteration=sweeps, CTM_Convergence_Epsilon=1e-30)
    Tn, _, n_ref, sig_ref = _oracle_run(43, 160, sweeps, 1e-30)
    genv = ta.Environment(gpu_ctx, 2, 2)
    genv.set_site_tensors(Tn)
    n, sig = genv.calc_ctm_environment(ta.peps_parameters(svd_power_iters=16, svd_warm_start=0, **pp))
    assert n == sweeps == n_ref
    assert abs(sig - sig_ref) <= 1e-8 * abs(sig_ref) + 1e-14, (sig, sig_ref)
