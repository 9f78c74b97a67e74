// GPU FP64 SVD used by the CTM projectors and the convergence check.
#pragma once

#include <functional>
#include <vector>

#include "tensor.hpp"

namespace tn {

struct SvdOptions {
  int power_iters = 4;        // subspace iterations for the range finder
  int max_sweeps = 30;        // Jacobi sweep cap
  double jacobi_tol = 0.0;    // 0: choose from the dimension
  // Optional start block of the range finder (n x p, p = 128 or 256 as the
  // range finder chooses for k, orthonormal columns,
  // e.g. a previous SvdResult::Vbasis); used when its shape fits, replacing
  // the Gaussian block, with start_power_iters iterations.
  const Tensor *start = nullptr;
  // Singular values belonging to the start columns (p values, descending);
  // columns with a value <= 1e-12 of the first are replaced by Gaussian ones.
  const Tensor *start_sigma = nullptr;
  int start_power_iters = 2;
  // Largest over smallest of those singular values (0: unknown).  Below
  // 1e7 the warm iterations of a 128-wide block (plain CholeskyQR passes;
  // a 256-wide block keeps the shifted ones) skip the orthonormalisation
  // between the A^T and the A product: (A A^T) then amplifies the rounding
  // left in a column towards the dominant directions by at most
  // range^2 * eps ~ 1e-2, which the next orthonormalisation removes.
  double start_range = 0.0;
};

struct SvdResult {
  Tensor U;                 // (m, k): left singular vectors as columns
  Tensor V;                 // (n, k): right singular vectors as columns
  Tensor s_dev;             // (k): singular values, descending
  std::vector<double> s;    // host copy of s_dev
  // host copy of all singular values of the projected problem (the p = 128
  // or 256 Ritz values on the randomized path), descending
  std::vector<double> s_all;
  // Randomized path only: all p right Ritz vectors (n x p), descending,
  // and their p singular values.
  Tensor Vbasis;
  Tensor sbasis;
};

// Leading k singular triplets of the m x n matrix A (row-major 2-D tensor).
// Exact one-sided Jacobi SVD when min(m, n) <= 128 (after a CholeskyQR
// reduction when max(m, n) > 128); otherwise a randomized subspace iteration
// with a p-wide block followed by the same Jacobi SVD of the projected
// problem, p = 128 for k <= 112 and 256 (blocked Cholesky, global-memory
// Jacobi) for 112 < k <= 240.  The returned k is min(k_target, rank-space
// size).
SvdResult svd_top(Context &ctx, const Tensor &A, int k_target,
                  const SvdOptions &opt);

// The same in two halves, for overlapping independent SVDs on the lanes of a
// forked Context: svd_top_begin issues all device work on the current stream
// (U, V and s_dev included) and an asynchronous copy of s into pinned host
// slot `slot` (distinct for SVDs in flight at the same time);
// svd_top_end waits for that stream and fills s.
struct SvdPending {
  SvdResult r;
  Context *ctx = nullptr;
  hipStream_t stream = nullptr;
  int slot = 0;
  int k = 0;
  int nall = 0;
};
SvdPending svd_top_begin(Context &ctx, const Tensor &A, int k_target,
                         const SvdOptions &opt, int slot);

// An m x n matrix known only through its products (matrix-free): mul(X) =
// A X for X (n x p), mulT(Y) = A^T Y for Y (m x p), both row-major 2-D
// tensors.  The randomized path calls only these; when min(m, n) <= 128 the
// matrix is formed from the products with the identity (n or m columns).
struct LinearOp {
  int64_t m = 0, n = 0;
  std::function<Tensor(const Tensor &)> mul, mulT;
};
SvdPending svd_top_begin(Context &ctx, const LinearOp &A, int k_target,
                         const SvdOptions &opt, int slot);
SvdResult svd_top_end(SvdPending &pending);
// The device results of a pending SVD without waiting for it (s and s_all
// stay empty; the values are in the pinned slot once the stream has passed
// the copy).
SvdResult svd_top_detach(SvdPending &pending);

// Singular values (descending, all min(m,n) of them) of a batch of equal
// shaped matrices, all of them <= 256 x 256.  Result: device tensor
// (batch, min(m,n)).
Tensor svd_values_batched(Context &ctx, const std::vector<Tensor> &mats);

// Orthonormal basis of the column space of Y (m x p row-major) by shifted
// CholeskyQR3 on the diagonally scaled Gram matrix.  If R != nullptr the
// p x p upper factor (Y = Q R) is returned.  passes = 2 (shifted + one plain
// pass) leaves every kept column normalised with cond(Q) ~ 1 + O(1e-2),
// enough for the intermediate blocks of a subspace iteration; passes = 1
// (shifted only) damps the directions below sqrt(shift) instead of
// normalising them.  shift_first = false makes every pass plain: for blocks
// whose columns are already nearly orthogonal (e.g. images of Ritz vectors),
// where the scaled Gram matrix is well conditioned.  Numerically dependent
// columns (relative pivot <= 1e-13 in the plain passes) come out as zero
// columns.
Tensor orthonormalize(Context &ctx, const Tensor &Y, Tensor *R, int passes = 3,
                      bool shift_first = true);

}  // namespace tn
