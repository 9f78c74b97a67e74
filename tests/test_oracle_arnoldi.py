"""Pins of the oracle's restarted Arnoldi (oracle/transfer.py: ArnoldiState,
arnoldi; TeNeS src/arnoldi.cpp:56-181).  No stored output of the reference
runs the restart (every test input keeps arnoldi_maxiterations = 1), so it is
pinned by what the mathematics fixes:

* a restart of p steps is p unshifted QR steps h <- q^T h q (identity() of
  tensor.hpp:58-64 returns a zero matrix, so the reference's shifts vanish):
  the new start vector is A^p Q[0], normalised (a transposed q, or a shifted
  step, fails this);
* a restart is a similarity of the Hessenberg block and keeps the basis
  orthonormal and the Arnoldi relation A Q_j = sum_i Q_i H[i, j] of the
  columns that precede the first disturbed one;
* an operator whose Krylov space closes within the basis converges in the
  first pass, where more passes change nothing;
* restarts drive the dominant eigenvalue to numpy's (known answer).
"""
import numpy as np
import pytest

from oracle import transfer as TMo


def _krylov_state(A, maxvec, seed=0):
    """A full pass of arnoldi.cpp's basis construction (no convergence stop)."""
    N = A.shape[0]
    st = TMo.ArnoldiState(N, maxvec, np.float64)
    st.Q[0] = np.random.default_rng(seed).standard_normal(N)
    st.orthonormalize(0)
    for k in range(1, maxvec + 1):
        st.Q[k] = A @ st.Q[k - 1]
        st.orthonormalize(k)
    return st


def _matrix(N=60, seed=1):
    rng = np.random.default_rng(seed)
    # nonsymmetric, one dominant real eigenvalue
    return rng.standard_normal((N, N)) / np.sqrt(N) + np.diag(np.r_[3.0, np.zeros(N - 1)])


@pytest.mark.parametrize("p", [1, 3, 6])
def test_restart_is_p_power_steps_on_the_start_vector(p):
    A = _matrix()
    k = 10
    st = _krylov_state(A, k)
    x = st.Q[0].copy()
    Qold = st.Q[:k].copy()
    h0 = st.H[:k, :k].copy()
    for _ in range(p):
        x = A @ x
    x /= np.linalg.norm(x)
    st.restart(k - p)
    assert abs(abs(st.Q[0] @ x) - 1.0) < 1e-12
    # the transposed turn of the basis (Q <- q Q) does not give the power step
    q, _ = np.linalg.qr(h0)
    assert abs(abs((q @ Qold)[0] @ (A @ Qold[0])) / np.linalg.norm(A @ Qold[0]) - 1.0) > 1e-3
    # nor does a step shifted by the largest Ritz value
    w = np.linalg.eigvals(h0)
    sigma = w[np.argmax(np.abs(w))].real
    qs, _ = np.linalg.qr(h0 - sigma * np.eye(k))
    y = A @ Qold[0]
    assert abs(abs((qs.T @ Qold)[0] @ y) / np.linalg.norm(y) - 1.0) > 1e-3


@pytest.mark.parametrize("minvec", [9, 6, 4])
def test_restart_is_a_similarity_and_keeps_the_arnoldi_relation(minvec):
    A = _matrix(seed=3)
    k = 12
    st = _krylov_state(A, k, seed=4)
    ev0 = np.sort_complex(np.linalg.eigvals(st.H[:k, :k]))
    st.restart(minvec)
    Q, H = st.Q, st.H
    # similarity (the 1e-12 cut touches only entries already at rounding level)
    ev1 = np.sort_complex(np.linalg.eigvals(H[:k, :k]))
    assert np.max(np.abs(ev1 - ev0)) < 1e-10 * np.max(np.abs(ev0))
    # orthonormal basis
    G = Q[:k] @ Q[:k].T
    assert np.max(np.abs(G - np.eye(k))) < 1e-12
    # the relation holds for the columns before the first disturbed one (the
    # p = k - minvec steps touch e_k^T q from position minvec - 1 on)
    for j in range(minvec - 1):
        r = A @ Q[j] - H[: j + 2, j] @ Q[: j + 2]
        assert np.max(np.abs(r)) < 1e-11, j
    # Hessenberg form survives: below the subdiagonal only the rounding of the
    # explicit steps remains (it grows with the number of shifts)
    assert np.max(np.abs(np.tril(H[:k, :k], -2))) < 1e-10 * np.max(np.abs(H[:k, :k]))


def test_closed_krylov_space_converges_in_the_first_pass():
    """Rank-6 operator, basis of 10: H[k, k-1] vanishes within the first
    pass, the residue drops below rtol, and 1 or 5 passes agree exactly."""
    rng = np.random.default_rng(7)
    U = np.linalg.qr(rng.standard_normal((80, 6)))[0]
    lam = np.array([2.0, -1.5, 1.0, 0.7, 0.4, 0.2])
    A = U @ np.diag(lam) @ U.T
    v0 = rng.standard_normal(80)
    one = TMo.arnoldi(lambda v: A @ v, v0, 3, 10, 5, 1, 1e-10)
    states = []
    five = TMo.arnoldi(lambda v: A @ v, v0, 3, 10, 5, 5, 1e-10, states)
    assert states == []
    assert np.array_equal(one, five)
    assert np.allclose(np.abs(one), [2.0, 1.5, 1.0], rtol=0, atol=1e-12)


def test_restarts_converge_to_the_dominant_eigenvalue():
    """Basis of 10, nev = 2, rtol 1e-14 (never reached): one pass leaves the
    dominant eigenvalue at 1.5e-4, each restart's 5 power steps improve it;
    after 4 passes it equals numpy's to 1e-10."""
    A = _matrix(N=120, seed=5)
    v0 = np.random.default_rng(6).standard_normal(120)
    w = np.linalg.eigvals(A)
    lead = w[np.argmax(np.abs(w))]
    err = []
    for maxiter in (1, 2, 4):
        states = []
        got = TMo.arnoldi(lambda v: A @ v, v0, 2, 10, 5, maxiter, 1e-14, states)
        assert len(states) == maxiter - 1
        err.append(abs(got[0] - lead) / abs(lead))
    assert err[0] > 1e-5 and err[1] < err[0] and err[2] < 1e-10, err


def test_more_eigenvalues_than_basis_vectors_is_refused():
    with pytest.raises(ValueError):
        TMo.arnoldi(lambda v: v, np.ones(30), 6, 5, 5, 1, 1e-10)
