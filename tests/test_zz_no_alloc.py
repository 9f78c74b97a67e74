"""tenes_svd_top_noalloc (the allocation-free variant of tenes_svd_top for
matrices of at most 128 x 128) through the C ABI, against tenes_svd_top on the
same inputs: two calls in a row, the second on a smaller matrix (the shared
scratch keeps the first call's contents), then the size limit."""
import ctypes as C

import numpy as np
import pytest

import tenes_amd as ta

pytestmark = pytest.mark.gpu


def _noalloc(ctx, A, m, n, k):
    kk = min(k, m, n)
    U = ta.core.Tensor_zeros(ctx, (m, kk))
    V = ta.core.Tensor_zeros(ctx, (n, kk))
    s = ta.core.Tensor_zeros(ctx, (kk,))
    kout = C.c_int(-1)
    ctx.check(ctx.lib.tenes_svd_top_noalloc(ctx.h, C.c_void_p(A.device_ptr()), m, n, k,
                                            C.c_void_p(U.device_ptr()), C.c_void_p(V.device_ptr()),
                                            C.c_void_p(s.device_ptr()), C.byref(kout)))
    ctx.synchronize()
    assert kout.value == kk
    return U.numpy(), s.numpy(), V.numpy()


@pytest.mark.parametrize("first,second", [((120, 100, 90), (37, 50, 20)), ((128, 128, 128), (9, 5, 5))])
def test_svd_top_noalloc_matches_svd_top(first, second):
    """Same kernels as tenes_svd_top's exact path on the same inputs: the
    singular values and vectors agree to rounding (the bar is 1e-13 relative
    to the largest singular value; the two calls run identical arithmetic);
    each result also reconstructs A and is orthonormal."""
    ctx = ta.Context(0, seed=5)
    rng = np.random.default_rng(11)
    for m, n, k in (first, second):
        A = rng.standard_normal((m, n)) * np.logspace(0, -6, n)  # graded columns
        At = ctx.tensor(A)
        U0, s0, V0 = ta.svd_top(At, k)
        U0, V0 = U0.numpy(), V0.numpy()
        U1, s1, V1 = _noalloc(ctx, At, m, n, k)
        smax = s0[0]
        assert s1.shape == s0.shape and U1.shape == U0.shape and V1.shape == V0.shape
        assert np.max(np.abs(s1 - s0)) <= 1e-13 * smax, (m, n, k)
        assert np.max(np.abs(U1 - U0)) <= 1e-12, (m, n, k)
        assert np.max(np.abs(V1 - V0)) <= 1e-12, (m, n, k)
        kk = min(m, n)
        if k >= kk:
            assert np.max(np.abs(U1 @ np.diag(s1) @ V1.T - A)) <= 1e-13 * smax * np.sqrt(m * n)
        assert np.max(np.abs(V1.T @ V1 - np.eye(V1.shape[1]))) < 1e-13
        ref = np.linalg.svd(A, compute_uv=False)[: s1.size]
        assert np.max(np.abs(s1 - ref)) < 1e-13 * smax


def test_svd_top_noalloc_rejects_large_matrices():
    ctx = ta.Context(0)
    A = ctx.tensor(np.ones((129, 4)))
    U = ta.core.Tensor_zeros(ctx, (129, 4))
    V = ta.core.Tensor_zeros(ctx, (4, 4))
    s = ta.core.Tensor_zeros(ctx, (4,))
    kout = C.c_int()
    rc = ctx.lib.tenes_svd_top_noalloc(ctx.h, C.c_void_p(A.device_ptr()), 129, 4, 4, C.c_void_p(U.device_ptr()),
                                       C.c_void_p(V.device_ptr()), C.c_void_p(s.device_ptr()), C.byref(kout))
    assert rc == -5
    assert "128 x 128" in ctx.lib.tenes_ctx_last_error(ctx.h).decode()
