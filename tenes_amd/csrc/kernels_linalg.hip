// FP64 dense linear algebra building blocks for the truncated-SVD projector.
//
//  * randn            : counter-based normal numbers for the range finder.
//  * chol_inv_upper   : shifted Cholesky G + shift I = R^T R and R^{-1}, one
//                       workgroup with the p x p matrix resident in LDS
//                       (p <= 128).  Used by shifted CholeskyQR; wider Gram
//                       matrices (p <= 256) are factorised as two blocks
//                       (chol_prep, chol_assemble).
//  * jacobi_svd       : one-sided (Hestenes) Jacobi SVD, one workgroup per
//                       matrix, the whole m x p matrix resident in LDS (in
//                       global memory above 128), the p/2 column pairs of a
//                       round-robin step processed by 64 groups of 16 lanes.  The (c, s) of every rotation
//                       is streamed to HBM so the accumulated right factor can
//                       be rebuilt row-parallel afterwards (jacobi_accumulate)
//                       without keeping a second p x p matrix in LDS.
//  * sort_desc        : rank sort of singular values, descending.
#include <cfloat>

#include "common.hpp"
#include "kernels.hpp"

namespace tn {
namespace kern {

namespace {

constexpr int kMaxDim = 128;
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void randn_kernel(double *p, int64_t n, uint64_t seed,
                             uint64_t counter) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint64_t j = counter + (uint64_t)(i >> 1);
    const uint64_t h1 = splitmix64(seed ^ (2 * j));
    const uint64_t h2 = splitmix64(seed ^ (2 * j + 1));
    const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;          // [0, 1)
    const double r = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    p[i] = (i & 1) ? r * sin(th) : r * cos(th);
  }
}

// One workgroup per column j of the n x p row-major block X: a column whose
// entries are all zero, or whose singular value sigma[j] is at most
// rel * sigma[0] (sigma may be null), is refilled with normal numbers (same
// generator and element numbering as randn_kernel).
__global__ void refill_columns_kernel(double *X, int64_t n, int p,
                                      const double *sigma, double rel,
                                      uint64_t seed, uint64_t counter) {
  __shared__ double part[256];
  const int j = blockIdx.x;
  const bool null_dir = sigma != nullptr && !(sigma[j] > rel * sigma[0]);
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = X[i * p + j];
    acc += v * v;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (part[0] != 0.0 && !null_dir) return;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const int64_t e = i * p + j;
    const uint64_t c = counter + (uint64_t)(e >> 1);
    const uint64_t h1 = splitmix64(seed ^ (2 * c));
    const uint64_t h2 = splitmix64(seed ^ (2 * c + 1));
    const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    X[e] = (e & 1) ? r * sin(th) : r * cos(th);
  }
}

// ---------------------------------------------------------------------------
// Shifted Cholesky + inverse of the upper factor: blocked, one workgroup of
// 4 waves, the padded matrix resident in LDS, diagonally scaled first.
//
// Blocks of 16.  Per block column: wave 0 factors the 16 x 16 diagonal block
// (lane l / register t holds element (l/16 + 4t, l%16); the pivot row is
// broadcast through a 16-entry LDS row, wave-synchronous, no barrier), the
// block row of R is a forward substitution with one thread per column, and
// the trailing update runs on the FP64 MFMA.  After the sweep the diagonal
// blocks are inverted in parallel (one 16-lane group per block) and the
// off-diagonal blocks of R^{-1} follow by block back substitution on the
// MFMA, whose output registers are directly the B operand of the next
// product.  Barriers: 3 per block column + 1 per block row of the inverse.
// ---------------------------------------------------------------------------
constexpr int kCholThreads = 256;
constexpr int kCholXdLd = 17;  // padded leading dimension of the diagonal inverses

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Orders one wave's LDS accesses around an exchange between its own lanes:
// the stores before the fence are complete before the loads after it.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

size_t chol_lds_bytes(int p) {
  const int pp = (p + 15) / 16 * 16;
  return ((size_t)pp * (pp + 1) + (size_t)(pp / 16) * 16 * kCholXdLd + 16) *
         sizeof(double);
}

__global__ __launch_bounds__(kCholThreads) void chol_inv_kernel(
    const double *G, int p, double shift_coef, double rel_floor, double *R,
    double *Rinv, const double *pre) {
  extern __shared__ double lds[];
  __shared__ int defic[kMaxDim];
  __shared__ double s_dg[kMaxDim], s_scl[kMaxDim];
  __shared__ double s_rd[kMaxDim];  // 1 / F[i][i] (0 for a dropped column)
  __shared__ double s_shift, s_thr, s_floor;
  __shared__ int s_zero, s_near;
  const int pp = (p + 15) / 16 * 16;  // padded size (identity padding)
  const int nb = pp / 16;
  const int ld = pp + 1;
  double *A = lds;                     // pp x ld, upper triangle: G, then F
  double *Xd = lds + (size_t)pp * ld;  // nb blocks of 16 x kCholXdLd: F_kk^-1
  double *rowbuf = Xd + (size_t)nb * 16 * kCholXdLd;  // 16: pivot row
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int lr = lane >> 4, lc = lane & 15;  // (row in 4-group, column)
  const int c0 = tid & 63, r0 = tid >> 6;    // load/store mapping

  // Load: thread -> columns c0, c0 + 64 and rows r0 + 4u; all 64 loads of a
  // thread are issued before the first LDS store (one memory round trip).
  {
    double v[32][2];
#pragma unroll
    for (int u = 0; u < 32; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = r0 + 4 * u, c = c0 + 64 * h;
        const bool in = r < p && c < p;
        const double g = G[in ? r * p + c : 0];
        v[u][h] = in ? g : (r == c ? 1.0 : 0.0);
      }
#pragma unroll
    for (int u = 0; u < 32; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = r0 + 4 * u, c = c0 + 64 * h;
        if (r < pp && c < pp) A[r * ld + c] = v[u][h];
      }
  }
  __syncthreads();
  // Diagonal (Jacobi) scaling G -> D G D, D = diag(G)^{-1/2}: the factor of
  // the scaled matrix is R~ with R = R~ D^{-1}, R^{-1} = D R~^{-1}.  Column
  // norms no longer enter the conditioning (a block of nearly orthogonal
  // columns of very different lengths factors as well as an orthonormal
  // one); zero columns keep a zero diagonal.
  // A pre-scaled input (pre != nullptr: a diagonal block or Schur
  // complement of a larger scaled and shifted Gram matrix) keeps unit scales.
  for (int i = tid; i < pp; i += kCholThreads) {
    const double g = A[i * ld + i];
    const double dg = (i < p && !pre) ? (g > 0.0 ? sqrt(g) : 0.0) : 1.0;
    s_dg[i] = dg;
    s_scl[i] = dg > 0.0 ? 1.0 / dg : 0.0;
  }
  __syncthreads();
  if (!pre) {
#pragma unroll 4
    for (int u = 0; u < 32; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = r0 + 4 * u, c = c0 + 64 * h;
        if (r < pp && c < pp) A[r * ld + c] = (A[r * ld + c] * s_scl[r]) * s_scl[c];
      }
    __syncthreads();
  }
  if (w == 0) {
    double tr = 0.0, mx = 0.0;
    for (int i = lane; i < p; i += 64) {
      const double g = A[i * ld + i];
      tr += fabs(g);
      mx = fmax(mx, g);
    }
    for (int o = 32; o > 0; o >>= 1) {
      tr += __shfl_xor(tr, o, 64);
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) {
      if (pre) {
        // shift already added; threshold and pivot floor of the whole matrix
        s_zero = 0;
        s_shift = 0.0;
        s_thr = pre[0];
        s_floor = pre[1];
      } else {
        s_zero = !(tr > 0.0);
        s_shift = shift_coef * tr;
        // pivots at or below rel_floor * max diagonal mark a column that is
        // numerically in the span of the previous ones (rank-deficient block)
        s_thr = rel_floor > 0.0 ? rel_floor * (mx + s_shift) : -1.0;
        // floor for pivots that rounding pushed to <= 0 when no deficiency
        // test is requested (first, shifted pass)
        s_floor = s_shift > 0.0 ? s_shift : DBL_MIN;
      }
    }
  }
  __syncthreads();
  if (s_zero) {
    // Zero (or NaN) Gram matrix: the input block is zero.  R = 0 (the block
    // has no range), Rinv = I so that Q = Y * Rinv stays zero.
    for (int r = r0; r < p; r += 4)
      for (int c = c0; c < p; c += 64) {
        if (R) R[r * p + c] = 0.0;
        Rinv[r * p + c] = (r == c) ? 1.0 : 0.0;
      }
    return;
  }
  // Near-identity Gram matrix (the second pass of a two-pass CholeskyQR):
  // with Gs = I + E, every |E_ij| <= 1e-10 and no shift, the factor is
  // R~ = I + U, U = triu(E) with halved diagonal, and R~^{-1} = I - U, both
  // exact to O(||U||^2) <= 2e-16 (||U||_2 <= 128 * 1e-10): no pivot chain.
  // No pivot can fall to the drop threshold (pivots are 1 +- 1e-10), so the
  // factorisation would drop nothing either.
  if (!pre && shift_coef == 0.0) {
    if (tid == 0) s_near = 1;
    __syncthreads();
    bool near = true;
#pragma unroll 4
    for (int u = 0; u < 32; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = r0 + 4 * u, c = c0 + 64 * h;
        if (r < p && c < p) near = near && fabs(A[r * ld + c] - (r == c ? 1.0 : 0.0)) <= 1e-10;
      }
    if (!near) s_near = 0;
    __syncthreads();
    if (s_near) {
#pragma unroll 4
      for (int u = 0; u < 32; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = r0 + 4 * u, c = c0 + 64 * h;
          if (r < p && c < p) {
            const double e = c > r ? A[r * ld + c] : (c == r ? 0.5 * (A[r * ld + r] - 1.0) : 0.0);
            const double id = r == c ? 1.0 : 0.0;
            if (R) R[r * p + c] = (id + e) * s_dg[c];
            Rinv[r * p + c] = s_scl[r] * (id - e);
          }
        }
      return;
    }
  }
  const double shift = s_shift, thr = s_thr;
  for (int i = tid; i < p; i += kCholThreads) A[i * ld + i] += shift;
  __syncthreads();
  const double floor_d = s_floor;

  for (int kb = 0; kb < nb; ++kb) {
    const int K0 = kb * 16;
    // ---- (1) diagonal block: Cholesky, wave 0 --------------------------
    if (w == 0) {
      double d[4], pv[4];
      bool dfc[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        d[t] = A[(K0 + lr + 4 * t) * ld + K0 + lc];
        pv[t] = 1.0;
        dfc[t] = false;
      }
      // right-looking, rows kept unscaled: D[i][j] -= D[k][i] D[k][j] / d_k
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (lr == (k & 3)) rowbuf[lc] = d[k >> 2];  // row k of the block
        wave_lds_fence();
        double piv = rowbuf[k];
        const bool dk = thr >= 0.0 && !(piv > thr);
        if (!dk && !(piv > floor_d)) piv = floor_d;
        const double rkc = dk ? 0.0 : rowbuf[lc] / piv;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int r = lr + 4 * t;
          if (r > k && lc >= r) d[t] -= rowbuf[r] * rkc;
          if (r == k) {
            pv[t] = piv;
            dfc[t] = dk;
          }
        }
        wave_lds_fence();  // all reads of this row before the next store
      }
      // F: scaled rows, unit pivot and zero row for dropped columns
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = lr + 4 * t;
        if (lc >= r) {
          const double sd = sqrt(pv[t]);
          A[(K0 + r) * ld + K0 + lc] =
              lc == r ? (dfc[t] ? 1.0 : sd) : (dfc[t] ? 0.0 : d[t] / sd);
          if (lc == r) s_rd[K0 + r] = dfc[t] ? 0.0 : 1.0 / sd;
        }
        if (lc == r) defic[K0 + r] = dfc[t] ? 1 : 0;
      }
    }
    __syncthreads();
    // ---- (2) block row of F by forward substitution, one thread per column:
    //   F_kk^T P = A[K0:K0+16, col], rows of dropped columns set to zero.
    // Column-oriented: once v_t is final (times 1/F[t][t], 0 for a dropped
    // column) it is eliminated from all later rows at once, so the chain
    // from one row to the next is one multiply and one FMA (the row-oriented
    // sums made every row wait for a chain of t FMAs and a division).
    const int ncol = pp - K0 - 16;
    if (tid < ncol) {
      const int col = K0 + 16 + tid;
      double v[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = A[(K0 + r) * ld + col];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        v[t] *= s_rd[K0 + t];
#pragma unroll
        for (int r = t + 1; r < 16; ++r) v[r] -= A[(K0 + t) * ld + K0 + r] * v[t];
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) A[(K0 + r) * ld + col] = v[r];
    }
    __syncthreads();
    // ---- (3) trailing update A[bi][bj] -= P[:,bi]^T P[:,bj], bi <= bj ------
    const int n = nb - kb - 1;
    const int nblk = n * (n + 1) / 2;
    for (int b = w; b < nblk; b += kCholThreads / 64) {
      int bi = 0, rem = b;
      while (rem >= n - bi) {  // row-wise enumeration of the upper blocks
        rem -= n - bi;
        ++bi;
      }
      const int I0 = (kb + 1 + bi) * 16, J0 = (kb + 1 + bi + rem) * 16;
      d4 acc;
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = A[(I0 + lr + 4 * q) * ld + J0 + lc];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double av = -A[(K0 + 4 * s + lr) * ld + I0 + lc];
        const double bv = A[(K0 + 4 * s + lr) * ld + J0 + lc];
        acc = mfma16(av, bv, acc);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) A[(I0 + lr + 4 * q) * ld + J0 + lc] = acc[q];
    }
    __syncthreads();
  }

  // ---- diagonal blocks of X = F^{-1}: 16-lane group g of wave w inverts
  // block 4w + g, lane lc holding column lc, rows from the bottom:
  //   X[i][c] = (delta_ic - sum_{k>i} F[i][k] X[k][c]) / F[i][i],
  // column-oriented as the block row above (x_i final, then eliminated from
  // the rows above it).  A dropped column i gets a zero row i here; its
  // column of X is zeroed on the store and no other column of X depends on
  // that row (row i of F is e_i^T, so the exact X[i][c] is 0 for c != i).
  {
    const int b = w * 4 + lr;
    if (b < nb) {
      const double *Fb = A + (size_t)b * 16 * ld + b * 16;
      double x[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = lc == i ? 1.0 : 0.0;
#pragma unroll
      for (int i = 15; i >= 0; --i) {
        x[i] *= s_rd[b * 16 + i];
#pragma unroll
        for (int r = 0; r < i; ++r) x[r] -= Fb[r * ld + i] * x[i];
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) Xd[(b * 16 + r) * kCholXdLd + lc] = x[r];
    }
  }
  __syncthreads();

  // ---- X = F^{-1}: block back substitution, block rows from the bottom ----
  // X[bi][bj] = -F_bi^{-1} sum_{bi<k<=bj} F[bi][k] X[k][bj]; off-diagonal
  // blocks of X are stored transposed in the (unused) lower triangle of A.
  // The k-sum uses two accumulators to halve the dependent MFMA chain.
  for (int bi = nb - 2; bi >= 0; --bi) {
    const int n = nb - 1 - bi;
    for (int idx = w; idx < n; idx += kCholThreads / 64) {
      const int bj = bi + 1 + idx;
      // T += F[bi][k] X[k][bj] for one block k
      auto accum = [&](d4 T, int k) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int kk = k * 16 + 4 * s + lr;
          const double av = A[(bi * 16 + lc) * ld + kk];
          const double bv = (k == bj) ? Xd[(bj * 16 + 4 * s + lr) * kCholXdLd + lc]
                                      : A[(bj * 16 + lc) * ld + kk];
          T = mfma16(av, bv, T);
        }
        return T;
      };
      d4 T0 = {0.0, 0.0, 0.0, 0.0}, T1 = {0.0, 0.0, 0.0, 0.0};
      int k = bi + 1;
      for (; k + 1 <= bj; k += 2) {
        T0 = accum(T0, k);
        T1 = accum(T1, k + 1);
      }
      if (k <= bj) T0 = accum(T0, k);
      const d4 T = T0 + T1;
      d4 Xo = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s)
        Xo = mfma16(Xd[(bi * 16 + lc) * kCholXdLd + 4 * s + lr], T[s], Xo);
#pragma unroll
      for (int q = 0; q < 4; ++q) A[(bj * 16 + lc) * ld + bi * 16 + lr + 4 * q] = -Xo[q];
    }
    __syncthreads();
  }

#pragma unroll 8
  for (int u = 0; u < 32; ++u)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 4 * u, c = c0 + 64 * h;
      if (r < p && c < p) {
        double rv = 0.0, xv = 0.0;
        if (c >= r) {
          rv = defic[r] ? 0.0 : A[r * ld + c];
          // a dropped column gets a zero column in R^{-1}: Q = Y R^{-1} then
          // has an exact zero there instead of an unnormalised residual (the
          // other columns of Q do not depend on it: row c of F is e_c^T)
          xv = defic[c] ? 0.0
                        : ((r >> 4) == (c >> 4) ? Xd[r * kCholXdLd + (c & 15)]
                                                : A[c * ld + r]);
        }
        if (R) R[r * p + c] = rv * s_dg[c];
        Rinv[r * p + c] = xv * s_scl[r];
      }
    }
}

// Cholesky of Gram matrices wider than one LDS-resident factorisation
// (128 < p <= 256), as two 128-blocks: with Gs = D G D + shift I (the
// scaled, shifted matrix of chol_inv_kernel) split at p1,
//   F11 = chol(Gs11), F12 = F11^{-T} Gs12, F22 = chol(Gs22 - F12^T F12),
//   F^{-1} = [X11, -X11 F12 X22; 0, X22]   (X = F^{-1} of each block),
// both block factorisations with the whole matrix's pivot threshold, so
// that a column of the second block that is dependent on the first one is
// dropped as the unblocked factorisation would drop it.
// chol_prep_kernel: scales (D, D^{-1}), shift, threshold and pivot floor,
// and the three blocks of Gs (contiguous, row-major).  A zero (or NaN) Gram
// matrix gives Gs = I with D = 0 and D^{-1} = 1: R = 0, R^{-1} = I.
__global__ __launch_bounds__(1024) void chol_prep_kernel(
    const double *G, int p, int p1, double coef, double rel_floor, double *G11,
    double *G12, double *G22, double *dg, double *scl, double *params) {
  __shared__ double s_scl[256];
  __shared__ double s_tr[16], s_mx[16];
  __shared__ double s_shift;
  __shared__ int s_zero;
  const int tid = threadIdx.x;
  const int p2 = p - p1;
  double tr = 0.0, mx = 0.0;
  for (int i = tid; i < p; i += blockDim.x) {
    const double g = G[(int64_t)i * p + i];
    const double d = g > 0.0 ? sqrt(g) : 0.0;
    const double sc = d > 0.0 ? 1.0 / d : 0.0;
    s_scl[i] = sc;
    const double gs = (g * sc) * sc;
    tr += fabs(gs);
    mx = fmax(mx, gs);
  }
  for (int o = 32; o > 0; o >>= 1) {
    tr += __shfl_xor(tr, o, 64);
    mx = fmax(mx, __shfl_xor(mx, o, 64));
  }
  if ((tid & 63) == 0) {
    s_tr[tid >> 6] = tr;
    s_mx[tid >> 6] = mx;
  }
  __syncthreads();
  if (tid == 0) {
    double t = 0.0, x = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      t += s_tr[w];
      x = fmax(x, s_mx[w]);
    }
    const int zero = !(t > 0.0);
    const double shift = coef * t;
    s_zero = zero;
    s_shift = zero ? 0.0 : shift;
    params[0] = rel_floor > 0.0 ? rel_floor * (x + shift) : -1.0;
    params[1] = shift > 0.0 ? shift : DBL_MIN;
  }
  __syncthreads();
  const int zero = s_zero;
  const double shift = s_shift;
  for (int i = tid; i < p; i += blockDim.x) {
    const double sc = s_scl[i];
    dg[i] = zero ? 0.0 : (sc > 0.0 ? 1.0 / sc : 0.0);
    scl[i] = zero ? 1.0 : sc;
  }
  for (int e = tid; e < p * p; e += blockDim.x) {
    const int r = e / p, c = e - (e / p) * p;
    const double v = zero ? (r == c ? 1.0 : 0.0)
                          : (G[e] * s_scl[r]) * s_scl[c] + (r == c ? shift : 0.0);
    if (r < p1) {
      if (c < p1)
        G11[r * p1 + c] = v;
      else
        G12[r * p2 + c - p1] = v;
    } else if (c >= p1) {
      G22[(r - p1) * p2 + c - p1] = v;
    }
  }
}

// F12 = F11^{-T} G12 by forward substitution (backward stable, unlike the
// product with the explicit inverse X11^T, whose error grows with cond(F11)
// and is large exactly where the block's columns are nearly dependent).  F11
// is the upper factor as chol_inv_upper writes it (a dropped column has a
// zero row, diagonal included; its row of F12 is zero).  One workgroup,
// F11 in LDS, one thread per column of G12 holding that column in
// registers, right-looking: v_t /= F[t][t], then v_r -= F[t][r] v_t for all
// r > t (independent updates, the F entries LDS broadcasts).
constexpr int kTrsmRows = 128;

__global__ __launch_bounds__(128) void trsm_ut_kernel(const double *F, int p2,
                                                      const double *G, double *V) {
  constexpr int n = kTrsmRows;
  extern __shared__ double Fs[];  // n x (n + 1)
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int r = e / n, c = e - (e / n) * n;
    Fs[r * (n + 1) + c] = F[e];
  }
  __syncthreads();
  if (tid >= p2) return;
  double v[n];
#pragma unroll
  for (int r = 0; r < n; ++r) v[r] = G[(int64_t)r * p2 + tid];
#pragma unroll
  for (int t = 0; t < n; ++t) {
    const double d = Fs[t * (n + 1) + t];
    v[t] = d != 0.0 ? v[t] / d : 0.0;
#pragma unroll
    for (int r = t + 1; r < n; ++r) v[r] -= Fs[t * (n + 1) + r] * v[t];
  }
#pragma unroll
  for (int r = 0; r < n; ++r) V[(int64_t)r * p2 + tid] = v[r];
}

// R and R^{-1} in the unscaled basis: R[r][c] = F[r][c] dg[c],
// R^{-1}[r][c] = scl[r] X[r][c], from the blocks (U = X11 F12 X22).
__global__ void chol_assemble_kernel(int p, int p1, const double *F11,
                                     const double *X11, const double *F12,
                                     const double *U, const double *F22,
                                     const double *X22, const double *dg,
                                     const double *scl, double *R, double *Rinv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= p * p) return;
  const int r = e / p, c = e - (e / p) * p;
  const int p2 = p - p1;
  double f = 0.0, x = 0.0;
  if (r < p1) {
    if (c < p1) {
      f = F11[r * p1 + c];
      x = X11[r * p1 + c];
    } else {
      f = F12[r * p2 + c - p1];
      x = -U[r * p2 + c - p1];
    }
  } else if (c >= p1) {
    f = F22[(r - p1) * p2 + c - p1];
    x = X22[(r - p1) * p2 + c - p1];
  }
  if (R) R[e] = f * dg[c];
  Rinv[e] = scl[r] * x;
}

// ---------------------------------------------------------------------------
// One-sided Jacobi SVD.
// ---------------------------------------------------------------------------
// Cross-lane moves inside a 16-lane row with DPP (register-to-register, no
// LDS round trip): quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror.  Each pairs every lane with one of the other half of the group
// being combined, so four steps give every lane the sum over its 16 lanes.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double sum16(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return v;
}

// Round-robin (circle method) pairing for pe (even) players, round s:
// pair 0 = (s, pe-1); pair k>0 = ((s+k) mod (pe-1), (s-k) mod (pe-1)).
__device__ __forceinline__ void rr_pair(int s, int k, int pe, int &i, int &j) {
  // 0 <= s, k < q: the residues mod q need one conditional subtraction each
  // (no integer division in the rotation loops)
  const int q = pe - 1;
  if (k == 0) {
    i = s;
    j = q;
  } else {
    i = s + k;
    if (i >= q) i -= q;
    j = s - k;
    if (j < 0) j += q;
  }
  if (i > j) {
    const int t = i;
    i = j;
    j = t;
  }
}

// FULL: m == kMaxDim (the randomized path's projected problems), so the
// row guards of the inner loops are compile-time true.
template <bool FULL>
__global__ __launch_bounds__(1024) void jacobi_kernel(
    double *Wg, int m_arg, int p, double tol, int max_sweeps, double *sigma,
    double *rot, int *nsteps) {
  const int m = FULL ? kMaxDim : m_arg;
  extern __shared__ double W[];  // pe columns of length m (col-major)
  __shared__ double s_off[64];
  __shared__ double s_tiny;
  const int b = blockIdx.x;
  double *Wb = Wg + (int64_t)b * m * p;
  const int pe = p + (p & 1);  // even number of columns (pad one zero col)
  const int tid = threadIdx.x;
  double fro = 0.0;
  const int nt = blockDim.x;
  for (int i0 = 0; i0 < m * pe; i0 += 8 * nt) {  // 8 loads in flight
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * nt + tid;
      v[u] = (i < m * p) ? Wb[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * nt + tid;
      if (i < m * pe) W[i] = v[u];
      fro += v[u] * v[u];
    }
  }
  for (int o = 32; o > 0; o >>= 1) fro += __shfl_xor(fro, o, 64);
  if ((tid & 63) == 0) s_off[tid >> 6] = fro;
  __syncthreads();
  if (tid == 0) {
    double f = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) f += s_off[w];
    // columns whose norm is below 1e-30 of the matrix norm are treated as
    // zero: neither rotated nor counted for convergence.  The bound is far
    // below rounding level on purpose: skipping a column of norm t leaves a
    // kept column of norm s unrotated by an angle of up to t / s, and kept
    // singular values go down to 1e-12 of the largest.
    s_tiny = f * 1e-60;
  }
  __syncthreads();
  const double tiny = s_tiny;
  const double tol2 = tol * tol;

  const int grp = tid >> 4;  // pair slot 0..63
  const int gl = tid & 15;   // lane in the 16-lane group
  const int npairs = pe / 2;
  const int64_t rot_stride = (int64_t)max_sweeps * (pe - 1) * npairs * 2;
  double *rotb = rot ? rot + (int64_t)b * rot_stride : nullptr;
  int steps = 0;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off_local = 0.0;
    for (int s = 0; s < pe - 1; ++s) {
      for (int k = grp; k < npairs; k += 64) {
        int i, j;
        rr_pair(s, k, pe, i, j);
        double *wi = W + i * m;
        double *wj = W + j * m;
        // the pair's two columns stay in registers between the inner
        // products and the rotation (m <= 128: 8 rows per lane)
        double xr[8], yr[8];
        double a = 0.0, bb = 0.0, g = 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int r = gl + 16 * u;
          xr[u] = (FULL || r < m) ? wi[r] : 0.0;
          yr[u] = (FULL || r < m) ? wj[r] : 0.0;
          a += xr[u] * xr[u];
          bb += yr[u] * yr[u];
          g += xr[u] * yr[u];
        }
        a = sum16(a);
        bb = sum16(bb);
        g = sum16(g);
        double c = 1.0, sn = 0.0;
        // rotate when |g| / sqrt(a bb) > tol, tested without a division
        const double g2 = g * g;
        if (g != 0.0 && a > tiny && bb > tiny && g2 > tol2 * (a * bb)) {
          {
            off_local = 1.0;
            // t = tan(theta) = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)) with
            // zeta = (bb - a) / (2 g), multiplied through by |2 g|
            const double dd = bb - a;
            const double t = (dd < 0.0 ? -2.0 * g : 2.0 * g) /
                             (fabs(dd) + sqrt(dd * dd + 4.0 * g2));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = c * t;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int r = gl + 16 * u;
              if (FULL || r < m) {
                wi[r] = c * xr[u] - sn * yr[u];
                wj[r] = sn * xr[u] + c * yr[u];
              }
            }
          }
        }
        if (rotb && gl == 0) {
          double *dst = rotb + ((int64_t)steps * npairs + k) * 2;
          dst[0] = c;
          dst[1] = sn;
        }
      }
      ++steps;
      __syncthreads();
    }
    // did any pair of this sweep exceed the tolerance (off_local = 1)?
    double v = off_local;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((tid & 63) == 0) s_off[tid >> 6] = v;
    __syncthreads();
    double off = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) off = fmax(off, s_off[w]);
    __syncthreads();
    if (!(off > 0.0)) break;
  }
  for (int jc = tid; jc < p; jc += blockDim.x) {
    double acc = 0.0;
    const double *wc = W + jc * m;
    for (int r = 0; r < m; ++r) acc += wc[r] * wc[r];
    sigma[(int64_t)b * p + jc] = sqrt(acc);
  }
  for (int i = tid; i < m * p; i += blockDim.x) Wb[i] = W[i];
  if (tid == 0 && nsteps) nsteps[b] = steps;
}

// One-sided Jacobi for matrices of up to kMaxDimBig columns and rows that
// do not fit in LDS: the same round-robin rounds, rotation formula,
// convergence test and rotation stream as jacobi_kernel, with the matrix
// in global memory (it stays in L2: 512 KB at 256 x 256).  One workgroup per
// matrix, 64 groups of 16 lanes, each group one column pair per step (the
// two columns in registers, 16 rows per lane); the workgroup barrier
// between rounds orders the rounds' global reads after the previous round's
// writes (all waves of a workgroup share one CU's cache).
constexpr int kMaxDimBig = 256;

__global__ __launch_bounds__(1024) void jacobi_big_kernel(
    double *Wg, int m, int p, double tol, int max_sweeps, double *sigma,
    double *rot, int *nsteps) {
  __shared__ double s_off[16];
  __shared__ double s_tiny;
  const int b = blockIdx.x;
  double *W = Wg + (int64_t)b * m * p;
  const int pe = p + (p & 1);
  const int tid = threadIdx.x;
  double fro = 0.0;
  for (int i = tid; i < m * p; i += blockDim.x) fro += W[i] * W[i];
  for (int o = 32; o > 0; o >>= 1) fro += __shfl_xor(fro, o, 64);
  if ((tid & 63) == 0) s_off[tid >> 6] = fro;
  __syncthreads();
  if (tid == 0) {
    double f = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) f += s_off[w];
    s_tiny = f * 1e-60;  // as in jacobi_kernel
  }
  __syncthreads();
  const double tiny = s_tiny;
  const double tol2 = tol * tol;
  const int grp = tid >> 4;
  const int gl = tid & 15;
  const int npairs = pe / 2;
  const int64_t rot_stride = (int64_t)max_sweeps * (pe - 1) * npairs * 2;
  double *rotb = rot ? rot + (int64_t)b * rot_stride : nullptr;
  int steps = 0;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off_local = 0.0;
    for (int s = 0; s < pe - 1; ++s) {
      for (int k = grp; k < npairs; k += 64) {
        int i, j;
        rr_pair(s, k, pe, i, j);
        // the padding column (p odd) is zero: its pairs never rotate
        const bool real_j = j < p;
        double *wi = W + (int64_t)i * m;
        double *wj = W + (int64_t)j * m;
        double xr[16], yr[16];
        double a = 0.0, bb = 0.0, g = 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int r = gl + 16 * u;
          xr[u] = r < m ? wi[r] : 0.0;
          yr[u] = (real_j && r < m) ? wj[r] : 0.0;
          a += xr[u] * xr[u];
          bb += yr[u] * yr[u];
          g += xr[u] * yr[u];
        }
        a = sum16(a);
        bb = sum16(bb);
        g = sum16(g);
        double c = 1.0, sn = 0.0;
        const double g2 = g * g;
        if (g != 0.0 && a > tiny && bb > tiny && g2 > tol2 * (a * bb)) {
          off_local = 1.0;
          const double dd = bb - a;
          const double t = (dd < 0.0 ? -2.0 * g : 2.0 * g) / (fabs(dd) + sqrt(dd * dd + 4.0 * g2));
          c = 1.0 / sqrt(1.0 + t * t);
          sn = c * t;
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int r = gl + 16 * u;
            if (r < m) {
              wi[r] = c * xr[u] - sn * yr[u];
              wj[r] = sn * xr[u] + c * yr[u];
            }
          }
        }
        if (rotb && gl == 0) {
          double *dst = rotb + ((int64_t)steps * npairs + k) * 2;
          dst[0] = c;
          dst[1] = sn;
        }
      }
      ++steps;
      __syncthreads();
    }
    double v = off_local;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((tid & 63) == 0) s_off[tid >> 6] = v;
    __syncthreads();
    double off = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) off = fmax(off, s_off[w]);
    __syncthreads();
    if (!(off > 0.0)) break;
  }
  for (int jc = tid; jc < p; jc += blockDim.x) {
    double acc = 0.0;
    const double *wc = W + (int64_t)jc * m;
    for (int r = 0; r < m; ++r) acc += wc[r] * wc[r];
    sigma[(int64_t)b * p + jc] = sqrt(acc);
  }
  if (tid == 0 && nsteps) nsteps[b] = steps;
}

// J = I * (product of recorded rotations).  Row-parallel: one 64-lane
// workgroup owns ROWS rows of J in LDS and replays the rotation stream.  The
// stream is read in chunks of kAccChunk rounds: the next chunk is loaded
// into registers (independent loads, all in flight) while the current one is
// replayed from LDS, so the replay does not wait on memory every round.
constexpr int kAccChunk = 16;

template <int ROWS, int MAXD>
__global__ __launch_bounds__(64) void jacobi_accum_kernel(const double *rot,
                                                          const int *nsteps,
                                                          int p, double *J) {
  __shared__ double row[ROWS][MAXD + 2];
  __shared__ double cs[kAccChunk * (MAXD / 2) * 2];
  const int pe = p + (p & 1);
  const int npairs = pe / 2;
  const int r0 = blockIdx.x * ROWS;
  const int lane = threadIdx.x;
  for (int q = 0; q < ROWS; ++q)
    for (int c = lane; c < pe; c += 64)
      row[q][c] = (r0 + q < p && c == r0 + q) ? 1.0 : 0.0;
  const int ns = nsteps[0];
  const int per = kAccChunk * npairs * 2;  // doubles per chunk (<= 2048)
  const int64_t total = (int64_t)ns * npairs * 2;
  const int nchunks = (ns + kAccChunk - 1) / kAccChunk;
  constexpr int kPerLane = kAccChunk * (MAXD / 2) * 2 / 64;
  double buf[kPerLane];
  auto fetch = [&](int ch) {
    const int64_t base = (int64_t)ch * per;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
      const int idx = u * 64 + lane;
      buf[u] = (idx < per && base + idx < total) ? rot[base + idx] : 0.0;
    }
  };
  if (nchunks > 0) fetch(0);
  int s = 0;  // round-robin step of round st: st mod (pe - 1)
  for (int ch = 0; ch < nchunks; ++ch) {
    __syncthreads();  // the previous chunk has been replayed
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
      const int idx = u * 64 + lane;
      if (idx < per) cs[idx] = buf[u];
    }
    __syncthreads();
    if (ch + 1 < nchunks) fetch(ch + 1);
    const int st0 = ch * kAccChunk;
    const int st1 = min(ns, st0 + kAccChunk);
    for (int st = st0; st < st1; ++st, s = (s + 1 == pe - 1) ? 0 : s + 1) {
      const double *cst = cs + (st - st0) * npairs * 2;
      for (int k = lane; k < npairs; k += 64) {
        int i, j;
        rr_pair(s, k, pe, i, j);
        const double c = cst[2 * k];
        const double sn = cst[2 * k + 1];
        // all rows' reads in flight before the first write (the rows are
        // distinct, so the writes cannot feed these reads)
        double x[ROWS], y[ROWS];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          x[q] = row[q][i];
          y[q] = row[q][j];
        }
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          row[q][i] = c * x[q] - sn * y[q];
          row[q][j] = sn * x[q] + c * y[q];
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  for (int q = 0; q < ROWS; ++q)
    if (r0 + q < p)
      for (int c = lane; c < p; c += 64) J[(int64_t)(r0 + q) * p + c] = row[q][c];
}

__global__ void sort_desc_kernel(const double *sigma, int p, int batch,
                                 double *sorted, int *perm) {
  const int b = blockIdx.x;
  const double *sv = sigma + (int64_t)b * p;
  for (int j = threadIdx.x; j < p; j += blockDim.x) {
    const double v = sv[j];
    int rank = 0;
    for (int k = 0; k < p; ++k) {
      const double u = sv[k];
      rank += (u > v) || (u == v && k < j);
    }
    sorted[(int64_t)b * p + rank] = v;
    perm[(int64_t)b * p + rank] = j;
  }
}

}  // namespace

int jacobi_max_dim() { return kMaxDimBig; }
int jacobi_lds_dim() { return kMaxDim; }

void linalg_init_device() {
  const int lds = kMaxDim * kMaxDim * (int)sizeof(double);
  TN_HIP_CHECK(hipFuncSetAttribute(
      reinterpret_cast<const void *>(&chol_inv_kernel),
      hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)chol_lds_bytes(kMaxDim)));
  TN_HIP_CHECK(hipFuncSetAttribute(
      reinterpret_cast<const void *>(&jacobi_kernel<true>),
      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  TN_HIP_CHECK(hipFuncSetAttribute(
      reinterpret_cast<const void *>(&jacobi_kernel<false>),
      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  TN_HIP_CHECK(hipFuncSetAttribute(
      reinterpret_cast<const void *>(&trsm_ut_kernel),
      hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)((size_t)kTrsmRows * (kTrsmRows + 1) * sizeof(double))));
}

void randn(double *p, int64_t n, uint64_t seed, uint64_t counter,
           hipStream_t s) {
  if (n <= 0) return;
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(randn_kernel, dim3((int)g), dim3(256), 0, s, p, n, seed,
                     counter);
  TN_HIP_CHECK(hipGetLastError());
}

void refill_columns(double *X, int64_t n, int p, const double *sigma,
                    double rel, uint64_t seed, uint64_t counter,
                    hipStream_t s) {
  if (n <= 0 || p <= 0) return;
  hipLaunchKernelGGL(refill_columns_kernel, dim3(p), dim3(256), 0, s, X, n, p,
                     sigma, rel, seed, counter);
  TN_HIP_CHECK(hipGetLastError());
}

void chol_inv_upper(const double *G, int p, double shift_coef,
                    double rel_floor, double *R, double *Rinv, hipStream_t s,
                    const double *pre) {
  TN_REQUIRE(p >= 1 && p <= kMaxDim, "chol_inv_upper: p out of range");
  const size_t lds = chol_lds_bytes(p);
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(kCholThreads), lds, s, G, p,
                     shift_coef, rel_floor, R, Rinv, pre);
  TN_HIP_CHECK(hipGetLastError());
}

int chol_block_dim() { return kMaxDim; }

void trsm_upper_t(const double *F, int p2, const double *G, double *V, hipStream_t s) {
  TN_REQUIRE(p2 >= 1 && p2 <= kTrsmRows, "trsm_upper_t: p2 out of range");
  const size_t lds = (size_t)kTrsmRows * (kTrsmRows + 1) * sizeof(double);
  hipLaunchKernelGGL(trsm_ut_kernel, dim3(1), dim3(128), lds, s, F, p2, G, V);
  TN_HIP_CHECK(hipGetLastError());
}

void chol_prep(const double *G, int p, int p1, double coef, double rel_floor,
               double *G11, double *G12, double *G22, double *dg, double *scl,
               double *params, hipStream_t s) {
  TN_REQUIRE(p > p1 && p1 >= 1 && p <= 256, "chol_prep: sizes out of range");
  hipLaunchKernelGGL(chol_prep_kernel, dim3(1), dim3(1024), 0, s, G, p, p1, coef,
                     rel_floor, G11, G12, G22, dg, scl, params);
  TN_HIP_CHECK(hipGetLastError());
}

void chol_assemble(int p, int p1, const double *F11, const double *X11,
                   const double *F12, const double *U, const double *F22,
                   const double *X22, const double *dg, const double *scl,
                   double *R, double *Rinv, hipStream_t s) {
  TN_REQUIRE(p > p1 && p1 >= 1 && p <= 256, "chol_assemble: sizes out of range");
  hipLaunchKernelGGL(chol_assemble_kernel, dim3(ceil_div(p * p, 256)), dim3(256), 0, s,
                     p, p1, F11, X11, F12, U, F22, X22, dg, scl, R, Rinv);
  TN_HIP_CHECK(hipGetLastError());
}

void jacobi_svd(double *W, int m, int p, int batch, double tol,
                int max_sweeps, double *sigma, double *rot, int *nsteps,
                hipStream_t s) {
  TN_REQUIRE(m >= 1 && p >= 1 && m <= kMaxDimBig && p <= kMaxDimBig,
             "jacobi_svd: dimension out of range");
  if (batch <= 0) return;
  if (m > kMaxDim || p > kMaxDim) {
    // above the LDS-resident size: the matrix stays in global memory
    hipLaunchKernelGGL(jacobi_big_kernel, dim3(batch), dim3(1024), 0, s, W, m, p, tol,
                       max_sweeps, sigma, rot, nsteps);
    TN_HIP_CHECK(hipGetLastError());
    return;
  }
  const int pe = p + (p & 1);
  const size_t lds = (size_t)m * pe * sizeof(double);
  if (m == kMaxDim)
    hipLaunchKernelGGL((jacobi_kernel<true>), dim3(batch), dim3(1024), lds, s, W, m, p,
                       tol, max_sweeps, sigma, rot, nsteps);
  else
    hipLaunchKernelGGL((jacobi_kernel<false>), dim3(batch), dim3(1024), lds, s, W, m, p,
                       tol, max_sweeps, sigma, rot, nsteps);
  TN_HIP_CHECK(hipGetLastError());
}

void jacobi_accumulate(const double *rot, const int *nsteps, int p, double *J,
                       hipStream_t s) {
  TN_REQUIRE(p >= 1 && p <= kMaxDimBig, "jacobi_accumulate: p out of range");
  // 2 rows per workgroup: measured 286 us per 128 x 128 replay at D=5,
  // chi=100 against 318 us with 4 rows and 394 us with 8
  constexpr int ROWS = 2;
  if (p <= kMaxDim)
    hipLaunchKernelGGL((jacobi_accum_kernel<ROWS, kMaxDim>), dim3(ceil_div(p, ROWS)),
                       dim3(64), 0, s, rot, nsteps, p, J);
  else
    hipLaunchKernelGGL((jacobi_accum_kernel<ROWS, kMaxDimBig>), dim3(ceil_div(p, ROWS)),
                       dim3(64), 0, s, rot, nsteps, p, J);
  TN_HIP_CHECK(hipGetLastError());
}

void sort_desc(const double *sigma, int p, int batch, double *sorted,
               int *perm, hipStream_t s) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(sort_desc_kernel, dim3(batch), dim3(128), 0, s, sigma, p,
                     batch, sorted, perm);
  TN_HIP_CHECK(hipGetLastError());
}

}  // namespace kern
}  // namespace tn
