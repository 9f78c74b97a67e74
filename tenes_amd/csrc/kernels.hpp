// Launch interfaces of the hand-written gfx950 kernels.  Host code in the
// library calls only these; every kernel runs on the stream it is given.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace tn {
namespace kern {

// ---------------------------------------------------------------------------
// Gather-GEMM:  C[m,n] = alpha * sum_k A[m,k] * B[k,n] + beta * C[m,n]
// with every operand addressed through offset tables, so any tensor
// permutation (tensordot leg order, output transpose) is fused into the
// tile loads and the epilogue store instead of a separate permute pass:
//   A[m,k] = A[a_m[m] + a_k[k]],  B[k,n] = B[b_k[k] + b_n[n]],
//   C[m,n] = C[c_m[m] + c_n[n]].
// a_mfast / b_nfast tell the loader which index walks contiguous memory.
// ---------------------------------------------------------------------------
struct GemmArgs {
  const double *A;
  const int *a_m;
  const int *a_k;
  const double *B;
  const int *b_k;
  const int *b_n;
  double *C;
  const int *c_m;
  const int *c_n;
  int M, N, K;
  double alpha, beta;
  bool a_mfast, b_nfast;
  // Affine fast path hints: when an offset table is an arithmetic progression
  // the loader computes the offset instead of reading the table.  0 = use the
  // table, otherwise the stride.
  int a_m_stride, a_k_stride, b_k_stride, b_n_stride;
  // Compute in FP32 on the MFMA (operands rounded to FP32 when they are
  // staged into LDS, FP32 accumulation; HBM data and the result stay FP64).
  bool fp32 = false;
  // Compute on the bf16 MFMA as three products of hi/lo bf16 halves of the
  // operands (split precision, FP32 accumulation; FP64 in HBM).
  bool bf16x3 = false;
};

// A GEMM launch shape: square output tile (64 or 128) and the number of K
// splits; bm = 0 selects the built-in heuristic.
struct GemmPlan {
  int bm = 0;
  int splits = 0;
};
// Launch shapes worth timing for (M, N, K), the heuristic's first.
// The stream-K launch of an M x N x K FP64 product (GemmPlan{128, 0}) exists:
// 128 x 128 tiles leaving at least half a round of 512 workgroups after the
// whole rounds, and K of at least 64 tiles of 16.
bool gemm_stream_k(int M, int N, int K);
std::vector<GemmPlan> gemm_candidates(int M, int N, int K);
// Scratch needed for split-K on the given shape (bytes, 0 if none).
size_t gemm_workspace_bytes(int M, int N, int K,
                            const GemmPlan *plan = nullptr);
// Returns the name of the kernel instantiation launched (static string).
// g.fp32 / g.bf16x3 select the FP32 or the split bf16 MFMA.
const char *gemm_f64(const GemmArgs &g, void *workspace, size_t ws_bytes,
                     hipStream_t s, const GemmPlan *plan = nullptr);
// The split bf16 variant with an explicit launch shape (square tile bm = 64
// or 128, K splits of k_per_split each; split-K partials in workspace).
const char *gemm_bf16x3(const GemmArgs &g, int bm, int splits, int k_per_split,
                        void *workspace, size_t ws_bytes, hipStream_t s);

// ---------------------------------------------------------------------------
// Double-layer site application (kernels_site.hip): for every block (s, f)
// of the environment legs of X (Ds x Df blocks, X block-major (s, f, kA,
// kB, bA, bB), all inner legs of dimension D),
//   out[(oA,oB),(boA,boB)] = sum TK[(kA,kB),p,(oA,oB)] TB[(bA,bB),(boA,boB),p]
//                                X[(kA,kB),(bA,bB)],
// written in the corner layout (oA, s, boA, oB, f, boB) or block-major
// (edge: s, f, oA, oB, boA, boB).  TK (kA kB, p, oA oB) and TB (bA bB,
// boA boB, p) contiguous.  Built for D = 2..5 and physical dimension 2.
// ---------------------------------------------------------------------------
struct SiteArgs {
  const double *X;
  const double *TK;
  const double *TB;
  double *out;
  int Ds, Df, D, p;
};
bool site_apply_supported(int D, int p);
void site_apply(const SiteArgs &a, bool edge, hipStream_t s);

// ---------------------------------------------------------------------------
// Elementwise / reduction helpers
// ---------------------------------------------------------------------------
void fill(double *p, int64_t n, double v, hipStream_t s);
// Pinned host slot (device address dst, 512 doubles): dst[0:n] = src[0:n]
// (n <= 511) and the int *count (if given) in the first 4 bytes of dst[511].
void write_slot(double *dst, const double *src, int n, const int *count,
                hipStream_t s);
// p (n x n, row-major) = identity
void eye(double *p, int64_t n, hipStream_t s);
// sc[i] = s[i]^(-1/2) while s[i] / s[0] > eps, else 0, for i < k (s on the
// device, descending): the projector truncation without a host round trip.
void projector_scale(const double *s_dev, int k, double eps, double *sc,
                     hipStream_t s);
// dst[dst_off(i)] = src[src_off(i)] for a rank<=8 row-major index space
// `shape` (used for transpose / slice / extend copies).
void strided_copy(const double *src, double *dst, int rank,
                  const int64_t *shape, const int64_t *src_strides,
                  const int64_t *dst_strides, int64_t n, hipStream_t s);
// out[0] = max_i |x_i|  (device scalar, two-pass, deterministic)
void max_abs(const double *x, int64_t n, double *out, double *partials,
             hipStream_t s);
int max_abs_partials(int64_t n);
// x[i] /= *d  (d is a device scalar)
void scale_by_inv(double *x, int64_t n, const double *d, hipStream_t s);
void scale(double *x, int64_t n, double a, hipStream_t s);
// out[0] = sum_i x[i]*y[i]
void dot(const double *x, const double *y, int64_t n, double *out,
         double *partials, hipStream_t s);
int dot_partials(int64_t n);
// y[i*inner + j + outer*...] *= v[(i / inner) % dim]  : scale along one axis
void scale_axis(double *x, int64_t n, const double *v, int64_t dim,
                int64_t inner, hipStream_t s);
// y += a * x
void axpy(double *y, const double *x, int64_t n, double a, hipStream_t s);
// y += (*a_dev) * x * sign
void axpy_dev(double *y, const double *x, int64_t n, const double *a_dev,
              double sign, hipStream_t s);
// out[r*k + q] = src[r*rs + perm[q]*cs] / (inv_scale ? scale[q] : 1)
// (a zero scale gives a zero column).  rows x k output, row-major.
void gather_cols(const double *src, int64_t rs, int64_t cs, const int *perm,
                 const double *scale, int rows, int k, double *out,
                 hipStream_t s);
// Split-K reduction: C[c_m[m]+c_n[n]] = alpha*sum_z P[z][m][n] + beta*C
void splitk_reduce(const double *P, int splits, int M, int N, double alpha,
                   double beta, double *C, const int *c_m, const int *c_n,
                   hipStream_t s);

// ---------------------------------------------------------------------------
// Dense linear algebra building blocks (FP64)
// ---------------------------------------------------------------------------
// Counter-based standard normal numbers (SplitMix64 hash + Box-Muller).
void randn(double *p, int64_t n, uint64_t seed, uint64_t counter,
           hipStream_t s);
// Columns of the n x p row-major X that are exactly zero, or whose
// singular value sigma[j] <= rel * sigma[0] (device array of p values, may
// be null), are refilled with normal numbers (seed, counter as randn).
void refill_columns(double *X, int64_t n, int p, const double *sigma,
                    double rel, uint64_t seed, uint64_t counter,
                    hipStream_t s);
// Shifted Cholesky of a symmetric p x p matrix G (row-major, p <= 128)
// using one workgroup: G + shift*I = R^T R, R upper triangular, shift =
// shift_coef * trace(G) (computed on the device), and Rinv = R^{-1}.
// R may be nullptr (not written).  With rel_floor > 0, a pivot <= rel_floor
// * max diagonal drops its column (rank-deficient input): that row of R is
// zero and that column of Rinv is zero, so Q = Y * Rinv has a zero column
// there (the other columns of Q do not depend on the dropped one).
// pre != nullptr: G is already scaled and shifted (a block of a larger
// factorisation); pre[0] is the pivot threshold (< 0: none), pre[1] the
// pivot floor, both device values written by chol_prep.
void chol_inv_upper(const double *G, int p, double shift_coef,
                    double rel_floor, double *R, double *Rinv, hipStream_t s,
                    const double *pre = nullptr);
// Largest p of one chol_inv_upper (the LDS-resident factorisation).
int chol_block_dim();
// The blocked factorisation for chol_block_dim() < p <= 256 (two blocks,
// split at p1): chol_prep scales and shifts G as chol_inv_upper does and
// writes its blocks G11 (p1 x p1), G12 (p1 x p2), G22 (p2 x p2), the scales
// dg = diag(G)^{1/2}, scl = 1/dg (0 for a zero column) and params (pivot
// threshold, pivot floor); chol_assemble writes R and R^{-1} (R may be
// nullptr) from the block factors F11, X11 = F11^{-1}, F12 = F11^{-T} G12,
// U = X11 F12 X22, F22, X22 = F22^{-1} of the pre-scaled blocks.
void chol_prep(const double *G, int p, int p1, double coef, double rel_floor,
               double *G11, double *G12, double *G22, double *dg, double *scl,
               double *params, hipStream_t s);
// V = F^{-T} G for the 128 x 128 upper factor F written by chol_inv_upper
// (rows of dropped columns zero: their rows of V are zero) and G 128 x p2
// (p2 <= 128), by forward substitution.
void trsm_upper_t(const double *F, int p2, const double *G, double *V, hipStream_t s);
void chol_assemble(int p, int p1, const double *F11, const double *X11,
                   const double *F12, const double *U, const double *F22,
                   const double *X22, const double *dg, const double *scl,
                   double *R, double *Rinv, hipStream_t s);
// One-sided (Hestenes) Jacobi SVD of a batch of p x p matrices with
// rows = m, one workgroup per matrix: held in LDS for m, p <= 128
// (jacobi_lds_dim), in global memory up to 256 (jacobi_max_dim).
// W is column-major per matrix (col j at W + j*m); on exit its columns are
// mutually orthogonal.  sigma[j] = ||W[:,j]||.  rotations (c,s) of every
// step are appended to rot (size max_sweeps*(p-1)*(p/2) pairs of doubles
// per matrix) unless rot == nullptr; nsteps[b] = number of recorded steps.
int jacobi_max_dim();
int jacobi_lds_dim();
void jacobi_svd(double *W, int m, int p, int batch, double tol,
                int max_sweeps, double *sigma, double *rot, int *nsteps,
                hipStream_t s);
// Rebuild the accumulated rotation matrix J (p x p, row-major) from the
// rotation stream written by jacobi_svd (batch of 1).
void jacobi_accumulate(const double *rot, const int *nsteps, int p, double *J,
                       hipStream_t s);
// sorted[b][r] = sigma[b][perm[b][r]], descending (ties keep index order).
void sort_desc(const double *sigma, int p, int batch, double *sorted,
               int *perm, hipStream_t s);
// One-time per-device kernel attribute setup (large dynamic LDS).
void linalg_init_device();

}  // namespace kern
}  // namespace tn
