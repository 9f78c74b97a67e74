"""GPU kernels against textbook references: the gather-GEMM behind every
tensordot (numpy tensordot/einsum in float64), transposes, and the SVD
(numpy's LAPACK SVD; reconstruction and orthogonality)."""
import numpy as np
import pytest

import tenes_amd as ta

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (1, 1, 1), (5, 7, 3), (64, 64, 64), (100, 37, 250),
                                   (300, 260, 129), (1024, 1024, 64), (3, 5, 5000), (250000 // 100, 50, 25)])
def test_matmul_shapes(gpu_ctx, M, N, K):
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C = ta.tensordot(gpu_ctx.tensor(A), gpu_ctx.tensor(B), [1], [0]).numpy()
    assert C.shape == (M, N)
    assert _rel(C, A @ B) < RTOL


@pytest.mark.parametrize("M,N,K", [(2500, 2500, 2500), (3500, 3500, 2000), (2100, 2100, 1024),
                                   (2000, 2300, 1500), (2500, 2500, 1100)])
def test_matmul_stream_k(gpu_ctx, M, N, K):
    """tenes_ctx_set_stream_k: row-major FP64 products whose 128 x 128 tile
    grid leaves a partial round (2500^2: 400 tiles on 512 workgroup slots)
    run stream-K: K
    iterations split evenly over 512 workgroups, split tiles summed by
    their last contributor.  Checked against numpy (1e-12 of max|C|), over
    three launches in a row (the arrival counters must come back to zero),
    bit-identical between launches (the shares are summed in rank order),
    and the launch must be the stream-K one (kernel report); (3500, 3500,
    2000) has a data-parallel round (512 tiles) before the split ones."""
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    ref = A @ B
    ta_, tb_ = gpu_ctx.tensor(A), gpu_ctx.tensor(B)
    gpu_ctx.set_profiling(True)
    gpu_ctx.set_stream_k(True)
    try:
        first = None
        for _ in range(3):
            C = ta.tensordot(ta_, tb_, [1], [0]).numpy()
            assert _rel(C, ref) < RTOL
            if first is None:
                first = C
            else:
                assert np.array_equal(C, first)
        names = [k for k in gpu_ctx.kernel_report() if f"[{M}x{N}x{K}]" in k]
    finally:
        gpu_ctx.set_profiling(False)
        gpu_ctx.set_stream_k(False)
    assert names and all("+sk" in k for k in names), names
    # and off (the default) the same product takes the data-parallel grid
    C = ta.tensordot(ta_, tb_, [1], [0]).numpy()
    assert _rel(C, ref) < RTOL


@pytest.mark.parametrize("case", [
    # (shapeA, shapeB, axA, axB, perm)
    ((4, 5, 6), (6, 5, 3), [1, 2], [1, 0], None),
    ((7, 3, 4, 2), (2, 4, 9), [2, 3], [1, 0], [2, 0, 1]),
    ((10, 10, 5, 5), (10, 5, 5, 11), [1, 2], [0, 1], [3, 1, 0, 2]),
    ((3, 4), (5, 6), [], [], None),
    ((20, 20, 4, 4), (20, 4, 4, 20), [0, 2, 3], [0, 1, 2], [1, 0]),
    ((2, 3, 2, 3, 2), (2, 3, 2, 3, 2), [0, 1, 4], [0, 1, 4], [1, 3, 0, 2]),
])
def test_tensordot_permuted(gpu_ctx, case):
    sa, sb, axa, axb, perm = case
    rng = np.random.default_rng(len(sa) * 7 + len(sb))
    A = rng.standard_normal(sa)
    B = rng.standard_normal(sb)
    ref = np.tensordot(A, B, axes=(axa, axb))
    if perm is not None:
        ref = ref.transpose(perm)
    got = ta.tensordot(gpu_ctx.tensor(A), gpu_ctx.tensor(B), axa, axb, perm).numpy()
    assert got.shape == ref.shape
    assert _rel(got, ref) < RTOL


@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (5, 7, 3), (300, 260, 129), (2500, 2500, 64), (64, 64, 5000)])
def test_matmul_fp32_compute(gpu_ctx, M, N, K):
    """The FP32-compute GEMM: inputs rounded to FP32, FP32 accumulation.
    Against the FP64 product the error is that of FP32 sums of K terms
    (bound: K * 2^-23 * sum|a||b|); its output layout is checked by the
    small shapes (a wrong row order would be an O(1) error)."""
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    gpu_ctx.set_fp32_contractions(True)
    try:
        C = ta.tensordot(gpu_ctx.tensor(A), gpu_ctx.tensor(B), [1], [0]).numpy()
    finally:
        gpu_ctx.set_fp32_contractions(False)
    ref = A @ B
    bound = K * 2.0 ** -23 * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(C - ref) <= bound), float(np.max(np.abs(C - ref) / bound))
    A32, B32 = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    assert _rel(C, A32 @ B32) < 1e-5
    # and the FP64 path is unaffected afterwards
    C64 = ta.tensordot(gpu_ctx.tensor(A), gpu_ctx.tensor(B), [1], [0]).numpy()
    assert _rel(C64, ref) < RTOL


def _bf16(x32):
    """float32 -> nearest-even bfloat16 value, as float32 (finite inputs)."""
    b = x32.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _split_bf16(x):
    x32 = x.astype(np.float32)
    hi = _bf16(x32)
    lo = _bf16((x32 - hi).astype(np.float32))
    return hi.astype(np.float64), lo.astype(np.float64)


@pytest.mark.parametrize("M,N,K,case", [(16, 16, 32, "nn"), (5, 7, 3, "nn"), (300, 260, 129, "nn"),
                                        (300, 260, 129, "tn"), (2500, 128, 2500, "tn"),
                                        (2500, 2500, 64, "nn"), (64, 64, 5000, "tt")])
def test_matmul_split_bf16(gpu_ctx, M, N, K, case):
    """The split-precision GEMM (contraction mode 2): both operands split
    into bf16 halves x = x_hi + x_lo (x_lo = bf16(x - x_hi), from the FP32
    value), three products A_hi B_hi + A_hi B_lo + A_lo B_hi on the bf16 MFMA
    with FP32 accumulation.  Against the same three products formed in FP64
    from the same halves (numpy) the difference is FP32 accumulation only
    (K * 2^-23 * sum|a||b|); against the FP64 product the split itself adds
    ~2^-16 per product.  Operand layouts "n"/"t" (row- or column-contiguous
    contracted index) exercise both loaders; a wrong MFMA lane layout would be
    an O(1) error at every shape."""
    rng = np.random.default_rng(M + 3 * N + 7 * K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    At = gpu_ctx.tensor(A.T.copy()) if case[0] == "t" else gpu_ctx.tensor(A)
    Bt = gpu_ctx.tensor(B.T.copy()) if case[1] == "t" else gpu_ctx.tensor(B)
    axa = [0] if case[0] == "t" else [1]
    axb = [1] if case[1] == "t" else [0]
    gpu_ctx.set_contraction_mode(2)
    try:
        C = ta.tensordot(At, Bt, axa, axb).numpy()
    finally:
        gpu_ctx.set_contraction_mode(0)
    Ah, Al = _split_bf16(A)
    Bh, Bl = _split_bf16(B)
    emu = Ah @ Bh + Ah @ Bl + Al @ Bh
    mag = np.abs(A) @ np.abs(B)
    assert np.all(np.abs(C - emu) <= (K + 2) * 2.0 ** -23 * mag), float(np.max(np.abs(C - emu) / mag))
    ref = A @ B
    err = np.max(np.abs(C - ref) / mag)
    assert err < 2.0 ** -14 + K * 2.0 ** -23, err
    # the FP64 path is unaffected afterwards
    C64 = ta.tensordot(At, Bt, axa, axb).numpy()
    assert _rel(C64, ref) < RTOL


def test_transpose(gpu_ctx):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 4, 5, 6, 2))
    for axes in [(4, 3, 2, 1, 0), (2, 3, 0, 1, 4), (1, 2, 3, 0, 4), (3, 0, 1, 2, 4)]:
        got = ta.transpose(gpu_ctx.tensor(A), axes).numpy()
        assert np.array_equal(got, A.transpose(axes))


def _check_svd(A, U, s, V, k, srel):
    s_ref = np.linalg.svd(A, compute_uv=False)[:k]
    assert np.max(np.abs(s - s_ref)) / s_ref[0] < srel
    # orthonormal columns and A V = U S on the kept triplets
    assert np.max(np.abs(U.T @ U - np.eye(k))) < 1e-10
    assert np.max(np.abs(V.T @ V - np.eye(k))) < 1e-10
    assert np.max(np.abs(A @ V - U * s)) / s[0] < 1e-10


@pytest.mark.parametrize("m,n", [(7, 7), (40, 40), (128, 128), (100, 60), (60, 100), (1, 5)])
def test_svd_direct_jacobi(gpu_ctx, m, n):
    rng = np.random.default_rng(m + 3 * n)
    A = rng.standard_normal((m, n))
    k = min(m, n)
    U, s, V = ta.svd_top(gpu_ctx.tensor(A), k)
    _check_svd(A, U.numpy(), s, V.numpy(), k, 1e-13)


@pytest.mark.parametrize("m,n", [(1000, 80), (80, 1000)])
def test_svd_qr_reduction(gpu_ctx, m, n):
    rng = np.random.default_rng(11)
    # graded columns: singular values spanning 12 decades
    Q1, _ = np.linalg.qr(rng.standard_normal((max(m, n), min(m, n))))
    Q2, _ = np.linalg.qr(rng.standard_normal((min(m, n), min(m, n))))
    sv = np.logspace(0, -12, min(m, n))
    A = (Q1 * sv) @ Q2.T
    if m < n:
        A = A.T
    k = 60
    U, s, V = ta.svd_top(gpu_ctx.tensor(A), k)
    _check_svd(A, U.numpy(), s, V.numpy(), k, 1e-13)


def test_svd_randomized_decaying(gpu_ctx):
    rng = np.random.default_rng(5)
    n = 1500
    Q1, _ = np.linalg.qr(rng.standard_normal((n, 400)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, 400)))
    sv = np.exp(-np.arange(400) / 12.0)
    A = (Q1 * sv) @ Q2.T
    k = 100
    U, s, V = ta.svd_top(gpu_ctx.tensor(A), k, power_iters=4)
    _check_svd(A, U.numpy(), s, V.numpy(), k, 1e-12)


def test_svd_rank_deficient(gpu_ctx):
    rng = np.random.default_rng(9)
    n = 600
    X = rng.standard_normal((n, 5))
    A = X @ rng.standard_normal((5, n))
    U, s, V = ta.svd_top(gpu_ctx.tensor(A), 20)
    s_ref = np.linalg.svd(A, compute_uv=False)[:20]
    assert np.max(np.abs(s[:5] - s_ref[:5])) / s_ref[0] < 1e-12
    assert np.all(s[5:] < 1e-12 * s[0])


def test_gemm_autotune():
    """tenes_ctx_set_autotune: the first contraction of a plan times every
    candidate launch shape and keeps the fastest; the tuning call and the
    tuned calls give the product to FP64 rounding (1e-12 relative), for a
    square GEMM, a long-K one (split-K candidates) and a permuted one."""
    import tenes_amd as ta

    ctx = ta.Context(0, seed=5)
    ctx.set_autotune(True)
    rng = np.random.default_rng(9)
    cases = [((600, 520), (520, 610), [1], [0]),
             ((96, 4000), (4000, 80), [1], [0]),
             ((40, 30, 50), (50, 30, 70), [1, 2], [1, 0])]
    for sa, sb, axa, axb in cases:
        A = rng.standard_normal(sa)
        B = rng.standard_normal(sb)
        ref = np.tensordot(A, B, (axa, axb))
        ta_, tb_ = ctx.tensor(A), ctx.tensor(B)
        for _ in range(3):
            C = ta.tensordot(ta_, tb_, axa, axb).numpy()
            assert np.max(np.abs(C - ref)) < 1e-12 * np.max(np.abs(ref)), (sa, sb)
    ctx.close()


def test_autotuned_ctm_matches_oracle():
    """A CTM run with autotuning on (fresh context): energy and Sz against the
    oracle to 1e-10, as without it."""
    import tenes_amd as ta
    from oracle import ctm as O, itps, models
    from oracle.lattice import SquareLattice
    from ctm_cases import su_tensors

    Tn, ham = su_tensors("heisenberg", 3, 60)
    Tn = list(Tn)
    lat = SquareLattice(2, 2)
    p = O.PEPSParameters(CHI=20, Max_CTM_Iteration=4, CTM_Convergence_Epsilon=0.0)
    env, _, _ = O.Calc_CTM_Environment(Tn, p, lat)
    ctx = ta.Context(0, seed=6)
    ctx.set_autotune(True)
    genv = ta.Environment(ctx, 2, 2)
    genv.set_site_tensors(Tn)
    genv.calc_ctm_environment(ta.peps_parameters(CHI=20, Max_CTM_Iteration=4, CTM_Convergence_Epsilon=0.0))
    E = 0.0
    gE = 0.0
    for s, leg in models.square_bonds(lat):
        dx, dy = (1, 0) if leg == 2 else (0, 1)
        gE += genv.contract_two_sites_op12(s, dx, dy, ham) / genv.contract_cluster(s, dx, dy)
    two = [{"name": "b", "group": 0, "source": s, "dx": 1 if leg == 2 else 0, "dy": 0 if leg == 2 else 1,
            "op": ham} for s, leg in models.square_bonds(lat)]
    w, _ = itps.measure_twosite(env, Tn, lat, two, [])
    E = np.real(sum(w[0].values()))
    assert abs(gE - E) < 1e-10 * abs(E), (gE, E)
    del genv
    ctx.close()
