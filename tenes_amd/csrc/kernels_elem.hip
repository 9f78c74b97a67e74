// Elementwise and reduction kernels (HBM-bound helpers) for gfx950.
// All reductions are two-pass with a fixed partial count so results are
// bitwise deterministic run to run.
#include "common.hpp"
#include "kernels.hpp"

namespace tn {
namespace kern {

namespace {

constexpr int TPB = 256;

__global__ void fill_kernel(double *p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// p (n x n) = identity
__global__ void eye_kernel(double *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n * n; i += stride) p[i] = (i / n == i % n) ? 1.0 : 0.0;
}

// dst[0:n] = src[0:n] and the int *count into the bit pattern of dst[511]:
// one dispatch writing a pinned host slot through its device address
// (instead of DMA copies).
__global__ void write_slot_kernel(double *dst, const double *src, int n,
                                  const int *count) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  if (threadIdx.x == 0 && count) {
    const int c = *count;
    reinterpret_cast<int *>(dst + 511)[0] = c;
  }
}

// Projector cut on the device (the truncation of ctm.cpp:275-288 for a
// descending spectrum): sc[i] = s[i]^(-1/2) while s[i] / s[0] > eps, else 0.
__global__ void projector_scale_kernel(const double *s, int k, double eps,
                                       double *sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const double s0 = s[0];
  sc[i] = (s0 > 0.0 && s[i] / s0 > eps) ? 1.0 / sqrt(s[i]) : 0.0;
}

struct CopyGeom {
  int rank;
  int64_t shape[8];
  int64_t ss[8];
  int64_t ds[8];
};

__global__ void strided_copy_kernel(const double *src, double *dst,
                                    CopyGeom cg, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    int64_t rem = i, so = 0, dof = 0;
    for (int d = cg.rank - 1; d >= 0; --d) {
      const int64_t idx = rem % cg.shape[d];
      rem /= cg.shape[d];
      so += idx * cg.ss[d];
      dof += idx * cg.ds[d];
    }
    dst[dof] = src[so];
  }
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <bool IS_MAX>
__device__ double block_reduce(double v) {
  __shared__ double sh[TPB / 64];
  v = IS_MAX ? wave_max(v) : wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = IS_MAX ? 0.0 : 0.0;
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int i = 1; i < TPB / 64; ++i) r = IS_MAX ? fmax(r, sh[i]) : r + sh[i];
  }
  return r;  // valid in thread 0 only
}

__global__ void max_abs_pass1(const double *x, int64_t n, double *part) {
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * TPB)
    v = fmax(v, fabs(x[i]));
  v = block_reduce<true>(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}
__global__ void max_pass2(const double *part, int np, double *out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += TPB) v = fmax(v, part[i]);
  v = block_reduce<true>(v);
  if (threadIdx.x == 0) out[0] = v;
}
__global__ void dot_pass1(const double *x, const double *y, int64_t n,
                          double *part) {
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * TPB)
    v += x[i] * y[i];
  v = block_reduce<false>(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}
__global__ void sum_pass2(const double *part, int np, double *out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += TPB) v += part[i];
  v = block_reduce<false>(v);
  if (threadIdx.x == 0) out[0] = v;
}

__global__ void scale_by_inv_kernel(double *x, int64_t n, const double *d) {
  const double inv = 1.0 / d[0];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] *= inv;
}
__global__ void scale_kernel(double *x, int64_t n, double a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] *= a;
}
__global__ void scale_axis_kernel(double *x, int64_t n, const double *v,
                                  int64_t dim, int64_t inner) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] *= v[(i / inner) % dim];
}
__global__ void axpy_kernel(double *y, const double *x, int64_t n, double a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] += a * x[i];
}
__global__ void axpy_dev_kernel(double *y, const double *x, int64_t n,
                                const double *a, double sign) {
  const double aa = sign * a[0];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] += aa * x[i];
}

__global__ void gather_cols_kernel(const double *src, int64_t rs, int64_t cs,
                                   const int *perm, const double *scale,
                                   int rows, int k, double *out) {
  const int64_t n = (int64_t)rows * k;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int r = (int)(i / k), q = (int)(i % k);
    double v = src[r * rs + (int64_t)perm[q] * cs];
    if (scale) {
      const double sc = scale[q];
      v = (sc > 0.0) ? v / sc : 0.0;
    }
    out[i] = v;
  }
}

inline int grid_for(int64_t n) {
  int64_t g = (n + TPB - 1) / TPB;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

void write_slot(double *dst, const double *src, int n, const int *count,
                hipStream_t s) {
  TN_REQUIRE(n >= 0 && n <= 511, "write_slot: at most 511 values");
  hipLaunchKernelGGL(write_slot_kernel, dim3(1), dim3(256), 0, s, dst, src, n,
                     count);
  TN_HIP_CHECK(hipGetLastError());
}

void eye(double *p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(eye_kernel, dim3(grid_for(n * n)), dim3(TPB), 0, s, p, n);
  TN_HIP_CHECK(hipGetLastError());
}

void projector_scale(const double *s_dev, int k, double eps, double *sc,
                     hipStream_t s) {
  if (k <= 0) return;
  hipLaunchKernelGGL(projector_scale_kernel, dim3((k + 255) / 256), dim3(256),
                     0, s, s_dev, k, eps, sc);
  TN_HIP_CHECK(hipGetLastError());
}

void fill(double *p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, p, n, v);
  TN_HIP_CHECK(hipGetLastError());
}

void strided_copy(const double *src, double *dst, int rank,
                  const int64_t *shape, const int64_t *src_strides,
                  const int64_t *dst_strides, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  TN_REQUIRE(rank >= 0 && rank <= 8, "strided_copy: rank > 8");
  CopyGeom cg;
  cg.rank = rank;
  for (int d = 0; d < rank; ++d) {
    cg.shape[d] = shape[d];
    cg.ss[d] = src_strides[d];
    cg.ds[d] = dst_strides[d];
  }
  hipLaunchKernelGGL(strided_copy_kernel, dim3(grid_for(n)), dim3(TPB), 0, s,
                     src, dst, cg, n);
  TN_HIP_CHECK(hipGetLastError());
}

int max_abs_partials(int64_t n) {
  int64_t g = (n + TPB * 8 - 1) / (TPB * 8);
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  return (int)g;
}
int dot_partials(int64_t n) { return max_abs_partials(n); }

void max_abs(const double *x, int64_t n, double *out, double *partials,
             hipStream_t s) {
  const int np = max_abs_partials(n);
  hipLaunchKernelGGL(max_abs_pass1, dim3(np), dim3(TPB), 0, s, x, n, partials);
  hipLaunchKernelGGL(max_pass2, dim3(1), dim3(TPB), 0, s, partials, np, out);
  TN_HIP_CHECK(hipGetLastError());
}

void dot(const double *x, const double *y, int64_t n, double *out,
         double *partials, hipStream_t s) {
  const int np = dot_partials(n);
  hipLaunchKernelGGL(dot_pass1, dim3(np), dim3(TPB), 0, s, x, y, n, partials);
  hipLaunchKernelGGL(sum_pass2, dim3(1), dim3(TPB), 0, s, partials, np, out);
  TN_HIP_CHECK(hipGetLastError());
}

void scale_by_inv(double *x, int64_t n, const double *d, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_by_inv_kernel, dim3(grid_for(n)), dim3(TPB), 0, s,
                     x, n, d);
  TN_HIP_CHECK(hipGetLastError());
}
void scale(double *x, int64_t n, double a, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, x, n, a);
  TN_HIP_CHECK(hipGetLastError());
}
void scale_axis(double *x, int64_t n, const double *v, int64_t dim,
                int64_t inner, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_axis_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, x,
                     n, v, dim, inner);
  TN_HIP_CHECK(hipGetLastError());
}
void gather_cols(const double *src, int64_t rs, int64_t cs, const int *perm,
                 const double *scale, int rows, int k, double *out,
                 hipStream_t s) {
  const int64_t n = (int64_t)rows * k;
  if (n <= 0) return;
  hipLaunchKernelGGL(gather_cols_kernel, dim3(grid_for(n)), dim3(TPB), 0, s,
                     src, rs, cs, perm, scale, rows, k, out);
  TN_HIP_CHECK(hipGetLastError());
}

void axpy(double *y, const double *x, int64_t n, double a, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, y, x, n,
                     a);
  TN_HIP_CHECK(hipGetLastError());
}
void axpy_dev(double *y, const double *x, int64_t n, const double *a_dev,
              double sign, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpy_dev_kernel, dim3(grid_for(n)), dim3(TPB), 0, s, y, x,
                     n, a_dev, sign);
  TN_HIP_CHECK(hipGetLastError());
}

}  // namespace kern
}  // namespace tn
