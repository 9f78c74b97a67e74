// Small dense FP64 tensors and factorizations on the host, for the
// full-update bond optimisation: its tensors hold at most ~1e4 entries (the
// reduced bond environment at D = 5, d = 2), where one kernel launch costs
// more than the arithmetic.  Leg conventions follow mptensor as TeNeS uses it:
// numpy-style transpose (new axis k = old axis perm[k]) and tensordot (free
// legs of A, then free legs of B).
#pragma once

#include <cstdint>
#include <vector>

namespace tn {
namespace host {

struct HT {
  std::vector<int64_t> shape;
  std::vector<double> v;  // row-major (last leg fastest)
  HT() = default;
  explicit HT(std::vector<int64_t> s);
  int rank() const { return (int)shape.size(); }
  int64_t size() const { return (int64_t)v.size(); }
};

HT transpose(const HT &t, const std::vector<int> &perm);
HT reshape(HT t, std::vector<int64_t> shape);
HT tensordot(const HT &A, const HT &B, const std::vector<int> &axA,
             const std::vector<int> &axB);
// Full contraction to a scalar: A leg axA[k] with B leg axB[k].
double trace(const HT &A, const HT &B, const std::vector<int> &axA,
             const std::vector<int> &axB);
HT add(const HT &a, const HT &b, double alpha, double beta);  // alpha a + beta b
// Scale along one leg by a vector (mptensor multiply_vector).
void scale_axis(HT &t, int axis, const std::vector<double> &s);
// Keep indices [0, n) of one leg.
HT take(const HT &t, int axis, int64_t n);

// Thin Householder QR of the m x n row-major matrix A: Q (m x k), R (k x n),
// k = min(m, n); Q has orthonormal columns even when A is rank deficient.
void qr(const std::vector<double> &A, int m, int n, std::vector<double> &Q,
        std::vector<double> &R);
// Symmetric eigendecomposition (cyclic Jacobi): w ascending, Z (n x n,
// row-major) with the eigenvectors as columns.
void eigh(const std::vector<double> &A, int n, std::vector<double> &w,
          std::vector<double> &Z);
// Thin SVD (one-sided Jacobi): U (m x k), s (k, descending), VT (k x n),
// k = min(m, n).  Columns of U belonging to zero singular values are zero.
void svd(const std::vector<double> &A, int m, int n, std::vector<double> &U,
         std::vector<double> &s, std::vector<double> &VT);

// Tensor forms: rows = row_legs, columns = col_legs (in the given order).
void qr(const HT &A, const std::vector<int> &row_legs,
        const std::vector<int> &col_legs, HT &Q, HT &R);
void svd(const HT &A, const std::vector<int> &row_legs,
         const std::vector<int> &col_legs, HT &U, std::vector<double> &s,
         HT &VT);

}  // namespace host
}  // namespace tn
