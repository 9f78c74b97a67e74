// GPU FP64 SVD for the CTM projectors: range finder + CholeskyQR + one-sided
// Jacobi, all on HBM-resident data, only the sorted singular values travel to
// the host (they decide the projector cut, as in TeNeS ctm.cpp:281-288).
#include "linalg.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "kernels.hpp"

namespace tn {

namespace {

constexpr int kJac = 128;  // LDS-resident Jacobi and Cholesky limit
constexpr int kJacBig = 256;  // global-memory Jacobi, blocked Cholesky limit
constexpr int kMinOversample = 16;  // range-finder columns beyond k

// Raw device buffer of ints carried by a Tensor (for lifetime management).
struct IntBuf {
  Tensor t;
  int *ptr = nullptr;
  IntBuf(Context &ctx, int64_t n) {
    t = empty(ctx, Shape{std::max<int64_t>((n + 1) / 2, 1)});
    ptr = reinterpret_cast<int *>(t.data());
  }
};

// Relative orthogonality |x.y| / (|x||y|) below which a column pair counts
// as converged: sqrt(m) * eps, the tolerance of LAPACK's dgesvj.
double jacobi_tol_for(const SvdOptions &opt, int m) {
  if (opt.jacobi_tol > 0.0) return opt.jacobi_tol;
  return std::sqrt(std::max(1.0, (double)m)) * 2.220446049250313e-16;
}

// Jacobi SVD of the rows x cols matrix held column-major in W (consumed).
// Returns sorted sigma (device, length cols), perm (device), J (cols x cols,
// row-major) if want_J.
struct JacobiOut {
  Tensor sigma_sorted;
  IntBuf *perm = nullptr;
  IntBuf *nsteps = nullptr;  // rounds taken (device int)
  Tensor J;
  Tensor W;
  ~JacobiOut() {
    delete perm;
    delete nsteps;
  }
};

void run_jacobi(Context &ctx, Tensor &W, int rows, int cols, bool want_J,
                const SvdOptions &opt, JacobiOut &out) {
  hipStream_t s = ctx.stream();
  const int pe = cols + (cols & 1);
  Tensor sigma = empty(ctx, Shape{cols});
  out.nsteps = new IntBuf(ctx, 1);
  int *nsteps = out.nsteps->ptr;
  Tensor rot;
  if (want_J)
    rot = empty(ctx, Shape{(int64_t)opt.max_sweeps * (pe - 1) * (pe / 2) * 2});
  kern::jacobi_svd(W.data(), rows, cols, 1, jacobi_tol_for(opt, rows),
                   opt.max_sweeps, sigma.data(), want_J ? rot.data() : nullptr,
                   nsteps, s);
  out.sigma_sorted = empty(ctx, Shape{cols});
  out.perm = new IntBuf(ctx, cols);
  kern::sort_desc(sigma.data(), cols, 1, out.sigma_sorted.data(),
                  out.perm->ptr, s);
  if (want_J) {
    out.J = empty(ctx, Shape{cols, cols});
    kern::jacobi_accumulate(rot.data(), nsteps, cols, out.J.data(), s);
  }
  out.W = W;
}

// Cholesky of the p x p Gram matrix G for 128 < p <= 256 as two blocks
// (kern::chol_prep): the same scaling, shift and pivot threshold as the
// one-workgroup factorisation, with the second block factorised from its
// Schur complement.  Writes R (if not null) and Ri = R^{-1}.
void chol_inv_blocked(Context &ctx, const Tensor &G, int p, double coef,
                      double rel_floor, Tensor *R, Tensor &Ri) {
  hipStream_t s = ctx.stream();
  const int p1 = kern::chol_block_dim(), p2 = p - p1;
  Tensor G11 = empty(ctx, Shape{p1, p1}), G12 = empty(ctx, Shape{p1, p2});
  Tensor G22 = empty(ctx, Shape{p2, p2});
  Tensor dg = empty(ctx, Shape{p}), scl = empty(ctx, Shape{p});
  Tensor prm = empty(ctx, Shape{2});
  kern::chol_prep(G.data(), p, p1, coef, rel_floor, G11.data(), G12.data(),
                  G22.data(), dg.data(), scl.data(), prm.data(), s);
  Tensor F11 = empty(ctx, Shape{p1, p1}), X11 = empty(ctx, Shape{p1, p1});
  kern::chol_inv_upper(G11.data(), p1, 0.0, 0.0, F11.data(), X11.data(), s,
                       prm.data());
  // F12 = F11^{-T} G12 by forward substitution (rows of dropped columns
  // zero, as in the one-block factorisation)
  Tensor F12 = empty(ctx, Shape{p1, p2});
  kern::trsm_upper_t(F11.data(), p2, G12.data(), F12.data(), s);
  Tensor S = tensordot(F12, F12, Axes{0}, Axes{0});    // p2 x p2
  kern::axpy(G22.data(), S.data(), (int64_t)p2 * p2, -1.0, s);  // Schur complement
  Tensor F22 = empty(ctx, Shape{p2, p2}), X22 = empty(ctx, Shape{p2, p2});
  kern::chol_inv_upper(G22.data(), p2, 0.0, 0.0, F22.data(), X22.data(), s,
                       prm.data());
  Tensor U = tensordot(tensordot(X11, F12, Axes{1}, Axes{0}), X22, Axes{1}, Axes{0});
  Ri = empty(ctx, Shape{p, p});
  kern::chol_assemble(p, p1, F11.data(), X11.data(), F12.data(), U.data(),
                      F22.data(), X22.data(), dg.data(), scl.data(),
                      R ? R->data() : nullptr, Ri.data(), s);
}

}  // namespace

Tensor orthonormalize(Context &ctx, const Tensor &Y_in, Tensor *R_out,
                      int passes, bool shift_first) {
  TN_REQUIRE(Y_in.rank() == 2, "orthonormalize: 2-D input expected");
  TN_REQUIRE(passes >= 1 && passes <= 3, "orthonormalize: 1 to 3 passes");
  const int64_t m = Y_in.shape[0], p = Y_in.shape[1];
  TN_REQUIRE(p >= 1 && p <= kJacBig, "orthonormalize: block width > 256");
  Tensor Y = Y_in;
  Tensor Rtot;
  const double u = 1.1102230246251565e-16;
  for (int pass = 0; pass < passes; ++pass) {
    Tensor G = tensordot(Y, Y, Axes{0}, Axes{0});  // p x p
    Tensor R = R_out ? empty(ctx, Shape{p, p}) : Tensor();
    Tensor Ri = empty(ctx, Shape{p, p});
    // shifted CholeskyQR3: the first pass is shifted (stable for any
    // condition number of Y), the next two are plain CholeskyQR with a
    // relative pivot floor that drops numerically dependent columns.  The
    // Gram matrix is diagonally scaled in the kernel, so plain passes are
    // enough when the columns are nearly orthogonal whatever their lengths.
    const bool shifted = pass == 0 && shift_first;
    const double coef =
        shifted ? 11.0 * (double)(m * p + p * (p + 1)) * u : 0.0;
    const double rel_floor = shifted ? 0.0 : 1e-13;
    if (p <= kJac)
      kern::chol_inv_upper(G.data(), (int)p, coef, rel_floor,
                           R_out ? R.data() : nullptr, Ri.data(), ctx.stream());
    else
      chol_inv_blocked(ctx, G, (int)p, coef, rel_floor, R_out ? &R : nullptr, Ri);
    Y = tensordot(Y, Ri, Axes{1}, Axes{0});
    if (R_out) Rtot = Rtot.defined() ? tensordot(R, Rtot, Axes{1}, Axes{0}) : R;
  }
  if (R_out) *R_out = Rtot;
  return Y;
}

namespace {

// The m x n matrix of a LinearOp with min(m, n) <= 128, from its products
// with the identity.
Tensor form_matrix(Context &ctx, const LinearOp &op) {
  if (op.n <= op.m) {
    Tensor I = empty(ctx, Shape{op.n, op.n});
    kern::eye(I.data(), op.n, ctx.stream());
    return op.mul(I);  // m x n
  }
  Tensor I = empty(ctx, Shape{op.m, op.m});
  kern::eye(I.data(), op.m, ctx.stream());
  return transpose(op.mulT(I), Axes{1, 0});  // (n x m)^T
}

SvdPending svd_top_begin_impl(Context &ctx, const Tensor *Ap, const LinearOp *op,
                              int k_target, const SvdOptions &opt, int slot);

}  // namespace

SvdPending svd_top_begin(Context &ctx, const LinearOp &op, int k_target,
                         const SvdOptions &opt, int slot) {
  TN_REQUIRE(op.m >= 1 && op.n >= 1 && op.mul && op.mulT, "svd_top: bad linear operator");
  if (std::min(op.m, op.n) <= kJac) {
    const Tensor A = form_matrix(ctx, op);
    return svd_top_begin_impl(ctx, &A, nullptr, k_target, opt, slot);
  }
  return svd_top_begin_impl(ctx, nullptr, &op, k_target, opt, slot);
}

SvdPending svd_top_begin(Context &ctx, const Tensor &A, int k_target,
                         const SvdOptions &opt, int slot) {
  return svd_top_begin_impl(ctx, &A, nullptr, k_target, opt, slot);
}

namespace {

SvdPending svd_top_begin_impl(Context &ctx, const Tensor *Ap, const LinearOp *op,
                              int k_target, const SvdOptions &opt, int slot) {
  TN_REQUIRE(slot >= 0 && slot < Context::kPinnedSlots, "svd_top_begin: bad pinned slot");
  ContractionPrecision fp64(ctx, false);  // the SVD and its QRs stay FP64
  // products with A: the explicit matrix, or the operator's (which keeps
  // the arithmetic its caller chose)
  const Tensor A = Ap ? *Ap : Tensor();
  auto prod = [&](const Tensor &X) { return Ap ? tensordot(A, X, Axes{1}, Axes{0}) : op->mul(X); };
  auto prodT = [&](const Tensor &Y) { return Ap ? tensordot(A, Y, Axes{0}, Axes{0}) : op->mulT(Y); };
  TN_REQUIRE(!Ap || A.rank() == 2, "svd_top: 2-D input expected");
  const int m = Ap ? (int)A.shape[0] : (int)op->m, n = Ap ? (int)A.shape[1] : (int)op->n;
  TN_REQUIRE(m >= 1 && n >= 1, "svd_top: empty matrix");
  TN_REQUIRE(k_target >= 1, "svd_top: k must be positive");
  if (k_target > kJacBig)
    throw Error("svd_top: at most 256 singular triplets are supported (" +
                    std::to_string(k_target) + " asked)",
                -5);
  hipStream_t s = ctx.stream();

  // A ~= Lq * X * Rq^T with orthonormal Lq (m x r), Rq (n x c); X is r x c.
  Tensor Lq, Rq, X;
  bool have_L = false, have_R = false, randomized = false;
  if (m <= kJac && n <= kJac) {
    X = A;
  } else if (n <= kJac) {
    Tensor Rf;
    Lq = orthonormalize(ctx, A, &Rf);  // A = Lq Rf, Rf n x n
    have_L = true;
    X = Rf;
  } else if (m <= kJac) {
    Tensor Rf;
    Tensor At = transpose(A, Axes{1, 0});  // n x m
    Rq = orthonormalize(ctx, At, &Rf);     // A^T = Rq Rf -> A = Rf^T Rq^T
    have_R = true;
    X = transpose(Rf, Axes{1, 0});
  } else {
    // range-finder block: 128 columns (LDS-resident Cholesky and Jacobi)
    // while that leaves 16 columns of oversampling, else 256 (blocked
    // Cholesky, global-memory Jacobi).  The subspace error after q
    // iterations is ~(s_{p+1}/s_k)^(2q+1): with fewer than 16 columns of
    // oversampling no iteration count bounds it.
    const int p = k_target <= kJac - kMinOversample ? kJac : kJacBig;
    if (k_target > p - kMinOversample)
      throw Error("svd_top: " + std::to_string(k_target) +
                      " leading triplets of a matrix wider than 128 need a "
                      "range-finder block wider than 256 (at most " +
                      std::to_string(p - kMinOversample) + " supported)",
                  -5);
    const bool warm = opt.start && opt.start->defined() && opt.start->rank() == 2 &&
                      opt.start->shape[0] == n && opt.start->shape[1] == p;
    Tensor Om;
    if (warm) {
      // Ritz directions of a numerically null singular value (an earlier,
      // lower-rank matrix) span nothing new and would keep the block in the
      // old range: give them fresh Gaussian directions
      Om = clone(*opt.start);
      const bool sv = opt.start_sigma && opt.start_sigma->defined() &&
                      opt.start_sigma->size() == p;
      kern::refill_columns(Om.data(), n, p, sv ? opt.start_sigma->data() : nullptr,
                           1e-12, ctx.seed,
                           ctx.next_rng_counter((uint64_t)Om.size()), s);
    } else {
      Om = empty(ctx, Shape{n, p});
      kern::randn(Om.data(), Om.size(), ctx.seed,
                  ctx.next_rng_counter((uint64_t)Om.size()), s);
    }
    // subspace iteration: the intermediate blocks only need every direction
    // normalised to a well-conditioned basis (shifted + one plain CholeskyQR
    // pass: cond ~ 1 + O(1e-2)); a single shifted pass is not enough, it
    // damps the directions below sqrt(shift) instead of normalising them and
    // the small kept singular directions die out over the iterations.  The
    // final range basis gets the full three passes.
    const int q = std::max(0, warm ? opt.start_power_iters : opt.power_iters);
    // A warm block holds Ritz vectors: while they stay aligned with the
    // singular directions (a settled environment), its images under A and
    // A^T are nearly orthogonal columns of very different lengths, which one
    // plain pass on the scaled Gram matrix orthonormalises (two for the
    // final bases); a Gaussian block mixes all directions and needs the
    // shifted passes.  The 256-wide block keeps the shifted passes warm too:
    // its Ritz values reach deeper, and with an environment still changing
    // (random D=2 tensors, chi = 160, 1e-2 per sweep) the misaligned images
    // lost directions to the plain pass's pivot floor (corner spectra off
    // by 1e-7 to 1e-3 against 2.5e-9 with shifted passes).
    const bool plain = warm && p <= kJac;
    const int mid = plain ? 1 : 2, fin = plain ? 2 : 3;
    Tensor Q = orthonormalize(ctx, prod(Om), nullptr, q == 0 ? fin : mid, !plain);
    // Skipping the orthonormalisation between the A^T and the A product
    // squares the conditioning of the next block: only for aligned warm
    // blocks.
    const bool skip_z = plain && opt.start_range > 0.0 && opt.start_range < 1e7;
    for (int it = 0; it < q; ++it) {
      Tensor Z = skip_z ? prodT(Q) : orthonormalize(ctx, prodT(Q), nullptr, mid, !plain);
      Q = orthonormalize(ctx, prod(Z), nullptr, it + 1 == q ? fin : mid, !plain);
    }
    Tensor R2;
    Tensor Bt = prodT(Q);  // n x p = A^T Q
    Rq = orthonormalize(ctx, Bt, &R2, fin, !plain);  // Bt = Rq R2
    Lq = Q;
    have_L = have_R = randomized = true;
    X = transpose(R2, Axes{1, 0});  // B = Q^T A = R2^T Rq^T
  }

  const int rows = (int)X.shape[0], cols = (int)X.shape[1];
  // Column-major copy of X = row-major X^T.
  Tensor W = transpose(X, Axes{1, 0});
  JacobiOut jo;
  run_jacobi(ctx, W, rows, cols, true, opt, jo);
  const int k = std::min(k_target, std::min(rows, cols));

  // Xhat = W(:,perm)/sigma (rows x k), Jp = J(:,perm) (cols x k)
  Tensor What = empty(ctx, Shape{rows, k});
  kern::gather_cols(jo.W.data(), 1, rows, jo.perm->ptr, jo.sigma_sorted.data(),
                    rows, k, What.data(), s);
  // the randomized path keeps all right Ritz vectors (a warm start block)
  const int kj = randomized ? cols : k;
  Tensor Jp = empty(ctx, Shape{cols, kj});
  kern::gather_cols(jo.J.data(), cols, 1, jo.perm->ptr, nullptr, cols, kj,
                    Jp.data(), s);
  SvdPending pd;
  SvdResult &r = pd.r;
  r.U = have_L ? tensordot(Lq, What, Axes{1}, Axes{0}) : What;
  if (randomized) {
    r.Vbasis = tensordot(Rq, Jp, Axes{1}, Axes{0});
    r.sbasis = jo.sigma_sorted;
    r.V = kj == k ? r.Vbasis : slice(r.Vbasis, 1, 0, k);
  } else {
    r.V = have_R ? tensordot(Rq, Jp, Axes{1}, Axes{0}) : Jp;
  }
  r.s_dev = slice(jo.sigma_sorted, 0, 0, k);
  // pinned slot: all <= 256 singular values, then the Jacobi round count,
  // written by one kernel through the slot's device address
  kern::write_slot(ctx.pinned_device(slot), jo.sigma_sorted.data(), cols,
                   jo.nsteps->ptr, s);
  pd.ctx = &ctx;
  pd.stream = s;
  pd.slot = slot;
  pd.k = k;
  pd.nall = cols;
  return pd;
}

}  // namespace

SvdResult svd_top_end(SvdPending &pd) {
  TN_REQUIRE(pd.ctx != nullptr, "svd_top_end: no pending SVD");
  TN_HIP_CHECK(hipStreamSynchronize(pd.stream));
  const double *host = pd.ctx->pinned(pd.slot);
  pd.r.s.assign(host, host + pd.k);
  pd.r.s_all.assign(host, host + pd.nall);
  int rounds = 0;
  std::memcpy(&rounds, host + 511, sizeof(int));
  pd.ctx->svd_calls++;
  pd.ctx->jacobi_rounds += rounds;
  pd.ctx = nullptr;
  return std::move(pd.r);
}

SvdResult svd_top_detach(SvdPending &pd) {
  TN_REQUIRE(pd.ctx != nullptr, "svd_top_detach: no pending SVD");
  pd.ctx->svd_calls++;
  pd.ctx = nullptr;
  return std::move(pd.r);
}

SvdResult svd_top(Context &ctx, const Tensor &A, int k_target,
                  const SvdOptions &opt) {
  SvdPending pd = svd_top_begin(ctx, A, k_target, opt, Context::kSlotSvdSync);
  return svd_top_end(pd);
}

Tensor svd_values_batched(Context &ctx, const std::vector<Tensor> &mats) {
  TN_REQUIRE(!mats.empty(), "svd_values_batched: empty batch");
  ContractionPrecision fp64(ctx, false);
  const int m = (int)mats[0].shape[0], n = (int)mats[0].shape[1];
  TN_REQUIRE(m <= kJacBig && n <= kJacBig,
             "svd_values_batched: matrices larger than 256 x 256");
  const int b = (int)mats.size();
  hipStream_t s = ctx.stream();
  // Jacobi acts on the columns of the stored column-major matrix.  Use the
  // shorter side as the column length: for m >= n store X column-major.
  const bool tr = m < n;  // then orthogonalise the columns of X^T
  const int rows = tr ? n : m, cols = tr ? m : n;
  Tensor W = empty(ctx, Shape{b, cols, rows});
  for (int i = 0; i < b; ++i) {
    TN_REQUIRE(mats[i].shape == mats[0].shape,
               "svd_values_batched: shape mismatch in batch");
    Tensor src = tr ? mats[i] : transpose(mats[i], Axes{1, 0});
    TN_HIP_CHECK(hipMemcpyAsync(W.data() + (int64_t)i * rows * cols,
                                src.data(), sizeof(double) * rows * cols,
                                hipMemcpyDeviceToDevice, s));
  }
  Tensor sigma = empty(ctx, Shape{b, cols});
  SvdOptions opt;
  kern::jacobi_svd(W.data(), rows, cols, b, jacobi_tol_for(opt, rows),
                   opt.max_sweeps, sigma.data(), nullptr, nullptr, s);
  Tensor sorted = empty(ctx, Shape{b, cols});
  IntBuf perm(ctx, (int64_t)b * cols);
  kern::sort_desc(sigma.data(), cols, b, sorted.data(), perm.ptr, s);
  return sorted;
}

}  // namespace tn
