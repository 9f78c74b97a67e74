// FP64 dense linear algebra building blocks for the truncated-SVD projector.
//
//  * randn            : counter-based normal numbers for the range finder.
//  * chol_inv_upper   : shifted Cholesky G + shift I = R^T R and R^{-1}, one
//                       workgroup with the p x p matrix resident in LDS
//                       (p <= 128).  Used by shifted CholeskyQR.
//  * jacobi_svd       : one-sided (Hestenes) Jacobi SVD, one workgroup per
//                       matrix, the whole m x p matrix resident in LDS, the
//                       p/2 column pairs of a round-robin step processed by
//                       64 groups of 16 lanes.  The (c, s) of every rotation
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

// ---------------------------------------------------------------------------
// Shifted Cholesky + inverse of the upper factor, one workgroup (1024 thr).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void chol_inv_kernel(const double *G, int p,
                                                        double shift_coef,
                                                        double rel_floor,
                                                        double *R,
                                                        double *Rinv) {
  extern __shared__ double A[];  // p*p, row-major
  __shared__ double s_shift, s_thr;
  __shared__ int s_zero;
  __shared__ int defic[kMaxDim];
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < p * p; i += nt) A[i] = G[i];
  __syncthreads();
  if (tid == 0) {
    double tr = 0.0, mx = 0.0;
    for (int i = 0; i < p; ++i) {
      tr += fabs(A[i * p + i]);
      mx = fmax(mx, A[i * p + i]);
    }
    s_zero = !(tr > 0.0);
    s_shift = shift_coef * tr;
    // pivots at or below rel_floor * max diagonal mark a column that is
    // numerically in the span of the previous ones (rank-deficient block)
    s_thr = rel_floor > 0.0 ? rel_floor * (mx + s_shift) : -1.0;
  }
  __syncthreads();
  if (s_zero) {
    // Zero (or NaN) Gram matrix: the input block is zero.  R = 0 (the block
    // has no range), Rinv = I so that Q = Y * Rinv stays zero.
    for (int i = tid; i < p * p; i += nt) {
      const double v = (i / p == i % p) ? 1.0 : 0.0;
      R[i] = 0.0;
      Rinv[i] = v;
    }
    return;
  }
  const double shift = s_shift, thr = s_thr;
  for (int i = tid; i < p; i += nt) A[i * p + i] += shift;
  __syncthreads();
  // Floor for pivots that rounding pushed to <= 0 when no deficiency test is
  // requested (first, shifted pass).
  const double floor_d = shift > 0.0 ? shift : DBL_MIN;
  for (int k = 0; k < p; ++k) {
    if (tid == 0) {
      double d = A[k * p + k];
      if (thr >= 0.0 && !(d > thr)) {
        // drop column k: unit pivot, zero row, no trailing update
        defic[k] = 1;
        A[k * p + k] = 1.0;
      } else {
        defic[k] = 0;
        if (!(d > floor_d)) d = floor_d;
        A[k * p + k] = sqrt(d);
      }
    }
    __syncthreads();
    const double rkk = A[k * p + k];
    const bool dk = defic[k] != 0;
    for (int j = k + 1 + tid; j < p; j += nt)
      A[k * p + j] = dk ? 0.0 : A[k * p + j] / rkk;
    __syncthreads();
    if (!dk) {
      const int rem = p - k - 1;
      for (int idx = tid; idx < rem * rem; idx += nt) {
        const int i = k + 1 + idx / rem;
        const int j = k + 1 + idx % rem;
        if (j >= i) A[i * p + j] -= A[k * p + i] * A[k * p + j];
      }
    }
    __syncthreads();
  }
  // R (upper) to HBM, zeros below the diagonal and in dropped rows.
  for (int i = tid; i < p * p; i += nt) {
    const int r = i / p, c = i % p;
    R[i] = (c >= r && !defic[r]) ? A[i] : 0.0;
  }
  __syncthreads();
  // In-place inversion of the upper factor, column by column (the leading
  // j x j block already holds its inverse):
  //   Rinv[0:j, j] = -(Rinv[0:j,0:j] * R[0:j, j]) / R[j][j],
  //   Rinv[j][j]   = 1 / R[j][j].
  // Eight lanes share one row i of the triangular mat-vec.
  __shared__ double y[kMaxDim];
  const int i_row = tid >> 3, part = tid & 7;
  for (int j = 0; j < p; ++j) {
    double acc = 0.0;
    if (i_row < j) {
      for (int k = i_row + part; k < j; k += 8)
        acc += A[i_row * p + k] * A[k * p + j];
    }
    acc += __shfl_xor(acc, 4, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 1, 64);
    if (i_row < j && part == 0) y[i_row] = acc;
    __syncthreads();
    const double inv_d = 1.0 / A[j * p + j];
    if (tid < j) A[tid * p + j] = -y[tid] * inv_d;
    __syncthreads();
    if (tid == 0) A[j * p + j] = inv_d;
    __syncthreads();
  }
  for (int i = tid; i < p * p; i += nt) {
    const int r = i / p, c = i % p;
    Rinv[i] = (c >= r) ? A[i] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// One-sided Jacobi SVD.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sum16(double v) {
  v += __shfl_xor(v, 8, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 1, 64);
  return v;
}

// Round-robin (circle method) pairing for pe (even) players, round s:
// pair 0 = (s, pe-1); pair k>0 = ((s+k) mod (pe-1), (s-k) mod (pe-1)).
__device__ __forceinline__ void rr_pair(int s, int k, int pe, int &i, int &j) {
  const int q = pe - 1;
  if (k == 0) {
    i = s;
    j = q;
  } else {
    i = (s + k) % q;
    j = (s - k + q) % q;
  }
  if (i > j) {
    const int t = i;
    i = j;
    j = t;
  }
}

__global__ __launch_bounds__(1024) void jacobi_kernel(
    double *Wg, int m, int p, double tol, int max_sweeps, double *sigma,
    double *rot, int *nsteps) {
  extern __shared__ double W[];  // pe columns of length m (col-major)
  __shared__ double s_off[64];
  const int b = blockIdx.x;
  double *Wb = Wg + (int64_t)b * m * p;
  const int pe = p + (p & 1);  // even number of columns (pad one zero col)
  const int tid = threadIdx.x;
  for (int i = tid; i < m * pe; i += blockDim.x)
    W[i] = (i < m * p) ? Wb[i] : 0.0;
  __syncthreads();

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
        double a = 0.0, bb = 0.0, g = 0.0;
        for (int r = gl; r < m; r += 16) {
          const double x = wi[r], y = wj[r];
          a += x * x;
          bb += y * y;
          g += x * y;
        }
        a = sum16(a);
        bb = sum16(bb);
        g = sum16(g);
        double c = 1.0, sn = 0.0;
        const double den = sqrt(a) * sqrt(bb);
        if (g != 0.0 && den > 0.0) {
          const double rel = fabs(g) / den;
          off_local = fmax(off_local, rel);
          if (rel > tol) {
            const double zeta = (bb - a) / (2.0 * g);
            const double t = copysign(1.0, zeta) /
                             (fabs(zeta) + hypot(1.0, zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = c * t;
            for (int r = gl; r < m; r += 16) {
              const double x = wi[r], y = wj[r];
              wi[r] = c * x - sn * y;
              wj[r] = sn * x + c * y;
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
    // workgroup max of the largest relative off-diagonal seen in this sweep
    double v = off_local;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((tid & 63) == 0) s_off[tid >> 6] = v;
    __syncthreads();
    double off = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) off = fmax(off, s_off[w]);
    __syncthreads();
    if (!(off > tol)) break;
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

// J = I * (product of recorded rotations).  Row-parallel: one 64-lane
// workgroup owns ROWS rows of J in LDS and replays the rotation stream.
template <int ROWS>
__global__ __launch_bounds__(64) void jacobi_accum_kernel(const double *rot,
                                                          const int *nsteps,
                                                          int p, double *J) {
  __shared__ double row[ROWS][kMaxDim + 2];
  const int pe = p + (p & 1);
  const int npairs = pe / 2;
  const int r0 = blockIdx.x * ROWS;
  const int lane = threadIdx.x;
  for (int q = 0; q < ROWS; ++q)
    for (int c = lane; c < pe; c += 64)
      row[q][c] = (r0 + q < p && c == r0 + q) ? 1.0 : 0.0;
  __syncthreads();
  const int ns = nsteps[0];
  for (int st = 0; st < ns; ++st) {
    const int s = st % (pe - 1);
    for (int k = lane; k < npairs; k += 64) {
      int i, j;
      rr_pair(s, k, pe, i, j);
      const double c = rot[((int64_t)st * npairs + k) * 2];
      const double sn = rot[((int64_t)st * npairs + k) * 2 + 1];
#pragma unroll
      for (int q = 0; q < ROWS; ++q) {
        const double x = row[q][i], y = row[q][j];
        row[q][i] = c * x - sn * y;
        row[q][j] = sn * x + c * y;
      }
    }
    __syncthreads();
  }
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

int jacobi_max_dim() { return kMaxDim; }

void linalg_init_device() {
  const int lds = kMaxDim * kMaxDim * (int)sizeof(double);
  TN_HIP_CHECK(hipFuncSetAttribute(
      reinterpret_cast<const void *>(&chol_inv_kernel),
      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  TN_HIP_CHECK(hipFuncSetAttribute(
      reinterpret_cast<const void *>(&jacobi_kernel),
      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
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

void chol_inv_upper(const double *G, int p, double shift_coef,
                    double rel_floor, double *R, double *Rinv, hipStream_t s) {
  TN_REQUIRE(p >= 1 && p <= kMaxDim, "chol_inv_upper: p out of range");
  const size_t lds = (size_t)p * p * sizeof(double);
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(1024), lds, s, G, p,
                     shift_coef, rel_floor, R, Rinv);
  TN_HIP_CHECK(hipGetLastError());
}

void jacobi_svd(double *W, int m, int p, int batch, double tol,
                int max_sweeps, double *sigma, double *rot, int *nsteps,
                hipStream_t s) {
  TN_REQUIRE(m >= 1 && p >= 1 && m <= kMaxDim && p <= kMaxDim,
             "jacobi_svd: dimension out of range");
  if (batch <= 0) return;
  const int pe = p + (p & 1);
  const size_t lds = (size_t)m * pe * sizeof(double);
  hipLaunchKernelGGL(jacobi_kernel, dim3(batch), dim3(1024), lds, s, W, m, p,
                     tol, max_sweeps, sigma, rot, nsteps);
  TN_HIP_CHECK(hipGetLastError());
}

void jacobi_accumulate(const double *rot, const int *nsteps, int p, double *J,
                       hipStream_t s) {
  TN_REQUIRE(p >= 1 && p <= kMaxDim, "jacobi_accumulate: p out of range");
  constexpr int ROWS = 4;
  hipLaunchKernelGGL((jacobi_accum_kernel<ROWS>), dim3(ceil_div(p, ROWS)),
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
