// The allocation-free variant of the small exact SVD (noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace tn {
namespace noalloc {

// Largest matrix side and the size of the per-device scratch array
// (bytes): work copy and accumulated rotations (2 x 128 x 128 doubles), the
// Jacobi rotation stream (30 sweeps x 127 rounds x 64 pairs x 2 doubles),
// singular values, permutation and round count.
constexpr int kSvdMaxDim = 128;
constexpr size_t kScratchBytes = 4166400;

// Looks up this device's address of the scratch array once (the device must
// be current); called when a context is created on the device.
void init_device(int device);

// Leading k singular triplets of the m x n row-major device matrix A
// (m, n <= 128) into caller-owned device buffers U (m x k'), V (n x k'),
// sv (k'), k' = min(k, m, n) returned in *k_out; work issued on stream s,
// no device allocation, no synchronisation.
void svd_small(int device, hipStream_t s, const double *A, int m, int n, int k,
               double *U, double *V, double *sv, int *k_out);

}  // namespace noalloc
}  // namespace tn
