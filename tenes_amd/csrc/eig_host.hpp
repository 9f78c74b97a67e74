// Small dense nonsymmetric eigenproblems on the host (n <= a few hundred):
// the Arnoldi Hessenberg matrix and the dense transfer matrix of small
// environments.  The role of LAPACK dgeev in TeNeS src/tensor.cpp:25-115.
#pragma once

#include <complex>
#include <vector>

namespace tn {

// Eigenvalues of the n x n real matrix A (row-major), sorted by |lambda|
// descending; the first `nev` are returned.  If last != nullptr, last[i] is
// the last component of the unit-2-norm right eigenvector of eigenvalue i.
void eig_nonsym(int n, const std::vector<double> &A, int nev,
                std::vector<std::complex<double>> &eigvals,
                std::vector<std::complex<double>> *last);

}  // namespace tn
