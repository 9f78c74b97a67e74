"""Transfer matrix of the CTM environment and Arnoldi restated in numpy
(TeNeS src/iTPS/transfer_matrix.cpp:84-551, src/arnoldi.cpp:34-181,
src/tensor.cpp:25-115)."""
from __future__ import annotations

import numpy as np

td = np.tensordot


class TMParameters:
    """transfer_matrix.hpp:32-47 defaults."""

    def __init__(self, **kw):
        self.num_eigvals = 4
        self.maxdim_dense_eigensolver = 200
        self.arnoldi_maxdim = 50
        self.arnoldi_restartdim = 20
        self.arnoldi_maxiter = 1
        self.arnoldi_rtol = 1e-10
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)


def eigen(A: np.ndarray, nev: int):
    """tensor.cpp:91-115: eigenvalues (and last components of the unit-norm
    right eigenvectors) sorted by |lambda| descending, first nev."""
    w, V = np.linalg.eig(A)
    order = np.argsort(-np.abs(w), kind="stable")[:nev]
    return w[order], V[-1, order]


class TransferMatrixCTM:
    def __init__(self, lat, Tn, env):
        self.lat = lat
        self.Tn = Tn
        if env.eTt[0].ndim == 3:
            # iTPO edges (chi, chi, D) (transfer_matrix.cpp:204-218, 399-409):
            # the same contractions with a bra leg of dimension one
            from .ctm import Env
            e4 = Env(len(env.C1))
            for k in ("C1", "C2", "C3", "C4"):
                setattr(e4, k, list(getattr(env, k)))
            for k in ("eTt", "eTr", "eTb", "eTl"):
                setattr(e4, k, [t[..., None] for t in getattr(env, k)])
            env = e4
        self.e = env

    def dim(self, d, fixed):
        lat, e = self.lat, self.e
        if d == 0:
            s0 = lat.index(0, fixed)
            s1 = lat.top(s0)
            return e.eTt[s0].shape[0] * e.eTb[s1].shape[1]
        s0 = lat.index(fixed, 0)
        s1 = lat.left(s0)
        return e.eTl[s0].shape[0] * e.eTr[s1].shape[1]

    def initial_vector(self, d, fixed):
        """transfer_matrix.cpp:162-185"""
        lat, e = self.lat, self.e
        if d == 0:
            site = lat.index(0, fixed)
            v = td(e.C1[site], e.C4[lat.top(site)], ([0], [1]))
        else:
            site = lat.index(fixed, 0)
            v = td(e.C4[site], e.C3[lat.left(site)], ([0], [1]))
        return v.reshape(-1)

    def matvec(self, d, vec, fixed):
        """transfer_matrix.cpp:187-308 (double-layer branch)."""
        lat, e = self.lat, self.e
        if d == 0:
            y = fixed
            s0 = lat.index(0, y)
            s1 = lat.top(s0)
            c0, c1 = e.eTt[s0].shape[0], e.eTb[s1].shape[1]
            out = vec.reshape(c0, c1)
            for x in range(lat.LX):
                site = lat.index(x, y)
                out = td(e.eTt[site], td(e.eTb[lat.top(site)], out, ([1], [1])), ([0, 2, 3], [3, 1, 2]))
            return out.reshape(-1)
        x = fixed
        x0 = x
        s0 = lat.index(x, 0)
        s1 = lat.left(s0)
        c0, c1 = e.eTl[s0].shape[0], e.eTr[s1].shape[1]
        out = vec.reshape(c0, c1)
        while True:
            for y in range(lat.LY):
                site = lat.index(x, y)
                out = td(e.eTl[site], td(e.eTr[lat.left(site)], out, ([1], [1])), ([0, 2, 3], [3, 1, 2]))
            x = (x + lat.skew + lat.LX) % lat.LX
            if x == x0:
                break
        return out.reshape(-1)

    def matrix(self, d, fixed):
        """transfer_matrix.cpp:379-475 (double-layer branch)."""
        lat, e = self.lat, self.e
        if d == 0:
            y = fixed
            s0 = lat.index(0, y)
            s1 = lat.top(s0)
            c0, c1 = e.eTt[s0].shape[0], e.eTb[s1].shape[1]
            res = td(e.eTt[s0], e.eTb[s1], ([2, 3], [2, 3])).transpose(0, 3, 1, 2)
            for x in range(1, lat.LX):
                site = lat.index(x, y)
                res = td(res, td(e.eTt[site], e.eTb[lat.top(site)], ([2, 3], [2, 3])), ([2, 3], [0, 3]))
            return res.reshape(c0 * c1, c0 * c1)
        x = fixed
        x0 = x
        s0 = lat.index(x, 0)
        s1 = lat.left(s0)
        c0, c1 = e.eTl[s0].shape[0], e.eTr[s1].shape[1]
        res = np.eye(c0 * c1, dtype=e.eTl[s0].dtype).reshape(c0, c1, c0, c1)
        while True:
            for y in range(lat.LY):
                site = lat.index(x, y)
                res = td(res, td(e.eTl[site], e.eTr[lat.left(site)], ([2, 3], [2, 3])), ([2, 3], [0, 3]))
            x = (x + lat.skew + lat.LX) % lat.LX
            if x == x0:
                break
        return res.reshape(c0 * c1, c0 * c1)

    def eigenvalues(self, d, fixed, prm: TMParameters):
        """transfer_matrix.cpp:84-142: dense eigensolver up to
        maxdim_dense_eigensolver, Arnoldi (with restarts) above."""
        N = self.dim(d, fixed)
        nev = min(prm.num_eigvals, N)
        if N == 1:
            return np.array([1.0 + 0j])
        if N <= prm.maxdim_dense_eigensolver:
            return eigen(self.matrix(d, fixed), nev)[0]
        maxvec, maxiter = prm.arnoldi_maxdim, prm.arnoldi_maxiter
        if N < maxvec:
            maxvec, maxiter = N, 1
        return arnoldi(lambda v: self.matvec(d, v, fixed), self.initial_vector(d, fixed), nev,
                       maxvec, prm.arnoldi_restartdim, maxiter, prm.arnoldi_rtol)


class ArnoldiState:
    """Krylov basis Q ((maxvec+1) x N) and Hessenberg matrix H ((maxvec+1) x
    maxvec) of arnoldi.cpp:34-41, one array each as there: a restart writes
    its maxvec x maxvec block of H and the later steps overwrite only the
    entries arnoldi.cpp's orthonormalize writes."""

    def __init__(self, N, maxvec, dtype):
        self.maxvec = maxvec
        self.Q = np.zeros((maxvec + 1, N), dtype=dtype)
        self.H = np.zeros((maxvec + 1, maxvec), dtype=dtype)

    def orthonormalize(self, k):
        """arnoldi.cpp:98-112: modified Gram-Schmidt of Q[k] against
        Q[0..k-1] (each projection taken from the already updated Q[k]),
        H[j, k-1] = <Q[j], Q[k]>, H[k, k-1] = ||Q[k]||."""
        Q, H = self.Q, self.H
        for j in range(k):
            x = np.vdot(Q[j], Q[k])
            H[j, k - 1] = x
            Q[k] = Q[k] - Q[j] * x
        x = np.sqrt(np.real(np.vdot(Q[k], Q[k])))
        Q[k] = Q[k] * (1.0 / x)
        if k > 0:
            H[k, k - 1] = x

    def residue(self, k, nev):
        """arnoldi.cpp:169-181: |H[k, k-1] * last component of the Ritz
        vector| / |Ritz value| of the nev largest Ritz values of H[:k, :k]."""
        w, last = eigen(self.H[:k, :k], nev)
        return np.abs(self.H[k, k - 1] * last) / np.abs(w)

    def restart(self, minvec, cutoff=1.0e-12):
        """arnoldi.cpp:117-167: maxvec - minvec QR steps h <- q^H h q of the
        Hessenberg block, each applied to the basis (Q[j] <- sum_i Q[i] q[i, j],
        j < maxvec); then entries of the block below cutoff in modulus are
        zeroed.

        Each step factors h - identity(k, sigma), sigma the i-th largest Ritz
        value.  identity (tensor.hpp:58-64) writes its value at index (k, k),
        outside the k x k matrix, in place of (i, i), so the matrix it returns
        is zero and every step is an unshifted QR step: q's first column is
        h e_0 / r_00 and the new start vector Q[0] is A Q[0] / r_00, a power
        step.  That is what the reference executes, and what this does (the
        Ritz values it computes for the shifts have no effect and are not
        computed here)."""
        k = self.maxvec
        h = self.H[:k, :k].copy()
        for _ in range(k - minvec):
            q, _ = np.linalg.qr(h)
            h = q.conj().T @ h @ q
            self.Q[:k] = q.T @ self.Q[:k]
        h[np.abs(h) ** 2 < cutoff * cutoff] = 0.0
        self.H[:k, :k] = h


def arnoldi(matvec, v0, nev, maxvec, mindim, maxiter, rtol, states=None):
    """arnoldi.cpp:43-96 (initialize + run) and :85-96 (eigenvalues): the nev
    eigenvalues of largest modulus of the operator matvec from the start
    vector v0, a Krylov basis of at most maxvec vectors and up to maxiter
    passes, restarted (ArnoldiState.restart) with mindim kept directions
    between passes.  `states`, when a list, receives a copy of (Q, H) after
    every restart."""
    N = v0.size
    if nev > maxvec:
        raise ValueError("Arnoldi: more eigenvalues than basis vectors")
    mindim = max(mindim, nev)
    st = ArnoldiState(N, maxvec, np.result_type(v0.dtype, np.float64))
    st.Q[0] = v0
    st.orthonormalize(0)
    for k in range(1, nev + 1):
        st.Q[k] = matvec(st.Q[k - 1])
        st.orthonormalize(k)
    res_nev = 1.0
    for it in range(1, maxiter + 1):
        for k in range(nev + 1, maxvec + 1):
            st.Q[k] = matvec(st.Q[k - 1])
            st.orthonormalize(k)
            res_nev = st.residue(k, nev)[nev - 1]
            if res_nev <= rtol:
                break
        if res_nev <= rtol or it == maxiter:
            break
        st.restart(mindim)
        if states is not None:
            states.append((st.Q.copy(), st.H.copy()))
    return eigen(st.H[:maxvec, :maxvec], nev)[0]


def correlation_length(eigvals, L):
    """correlation_length.cpp:97-124: xi = -L / log|t1/t0| and
    e_i = -log|t_i/t0| / L."""
    if len(eigvals) == 1:
        return 0.0, []
    e0 = abs(eigvals[0])
    xi = -L / np.log(abs(eigvals[1]) / e0)
    return xi, [-np.log(abs(v) / e0) / L for v in eigvals[1:]]
