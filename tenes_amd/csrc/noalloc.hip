// The small exact SVD of tenes_svd_top (both dimensions <= 128: one-sided
// Jacobi on the whole matrix) without any device allocation, for callers
// that cannot allocate device memory.  Its scratch -- the column-major work
// copy, the Jacobi rotation stream, the singular values before and after
// sorting, the permutation, the round count and the accumulated rotations --
// lives in one file-scope device array of fixed size.  Every region is
// written before it is read within a call, so nothing needs clearing between
// calls; calls that overlap on one device share the array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "common.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"

namespace tn {
namespace noalloc {

namespace {

constexpr int kDim = kSvdMaxDim;  // 128
constexpr int kSweeps = 30;       // Jacobi sweep cap (SvdOptions default)
constexpr size_t kAlign = 32;     // doubles (256 bytes)

constexpr size_t up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// region offsets in doubles
constexpr size_t kW = 0;                                     // m x n
constexpr size_t kJ = kW + up((size_t)kDim * kDim);          // n x n
constexpr size_t kSigma = kJ + up((size_t)kDim * kDim);      // n
constexpr size_t kSorted = kSigma + up(kDim);                // n
constexpr size_t kPerm = kSorted + up(kDim);                 // n ints
constexpr size_t kSteps = kPerm + up(kDim / 2);              // 1 int
constexpr size_t kRot = kSteps + up(1);                      // rotation stream
constexpr size_t kRotLen = (size_t)kSweeps * (kDim - 1) * (kDim / 2) * 2;
constexpr size_t kTotal = kRot + up(kRotLen);

static_assert(kTotal * sizeof(double) == kScratchBytes,
              "scratch size constant out of date");

}  // namespace

}  // namespace noalloc
}  // namespace tn

// File-scope device array holding every call's scratch (per device).  Host
// code never passes its name to a kernel: the host symbol is a stand-in, and
// each device's address comes from hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[tn::noalloc::kTotal];

namespace tn {
namespace noalloc {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

}  // namespace

void init_device(int device) {
  TN_REQUIRE(device >= 0 && device < kMaxDevices, "no-alloc scratch: device out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return;
  int cur = -1;
  TN_HIP_CHECK(hipGetDevice(&cur));
  TN_REQUIRE(cur == device, "no-alloc scratch: device must be current at set-up");
  void *p = nullptr;
  TN_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  TN_REQUIRE(p != nullptr, "no-alloc scratch: no device address");
  g_ptr[device] = static_cast<double *>(p);
}

void svd_small(int device, hipStream_t s, const double *A, int m, int n, int k,
               double *U, double *V, double *sv, int *k_out) {
  if (m < 1 || n < 1 || m > kDim || n > kDim)
    throw Error("tenes_svd_top_noalloc: the matrix must be between 1 x 1 and " +
                    std::to_string(kDim) + " x " + std::to_string(kDim) + " (got " +
                    std::to_string(m) + " x " + std::to_string(n) + ")",
                -5);
  TN_REQUIRE(k >= 1, "tenes_svd_top_noalloc: k must be positive");
  TN_REQUIRE(A && U && V && sv, "tenes_svd_top_noalloc: null device pointer");
  TN_REQUIRE(device >= 0 && device < kMaxDevices, "tenes_svd_top_noalloc: bad device");
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = g_ptr[device];
  }
  TN_REQUIRE(base != nullptr,
             "tenes_svd_top_noalloc: scratch of this device not set up (create a "
             "context on it first)");
  double *W = base + kW;
  double *J = base + kJ;
  double *sigma = base + kSigma;
  double *sorted = base + kSorted;
  int *perm = reinterpret_cast<int *>(base + kPerm);
  int *nsteps = reinterpret_cast<int *>(base + kSteps);
  double *rot = base + kRot;
  // W = A^T row-major = A column-major: the Jacobi kernel orthogonalises the
  // n columns (length m) of A, as svd_top does on this path
  const int64_t shape[2] = {n, m}, src[2] = {1, n}, dst[2] = {m, 1};
  kern::strided_copy(A, W, 2, shape, src, dst, (int64_t)m * n, s);
  const double tol = std::sqrt((double)m) * 2.220446049250313e-16;
  kern::jacobi_svd(W, m, n, 1, tol, kSweeps, sigma, rot, nsteps, s);
  kern::sort_desc(sigma, n, 1, sorted, perm, s);
  kern::jacobi_accumulate(rot, nsteps, n, J, s);
  const int kk = std::min(k, std::min(m, n));
  // U = W(:, perm) / sigma (m x kk), V = J(:, perm) (n x kk), s = sorted[:kk]
  kern::gather_cols(W, 1, m, perm, sorted, m, kk, U, s);
  kern::gather_cols(J, n, 1, perm, nullptr, n, kk, V, s);
  TN_HIP_CHECK(hipMemcpyAsync(sv, sorted, (size_t)kk * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
  if (k_out) *k_out = kk;
}

}  // namespace noalloc
}  // namespace tn
