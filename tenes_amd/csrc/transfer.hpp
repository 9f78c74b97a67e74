// Row/column transfer matrix of the CTM environment and its leading
// eigenvalues (TeNeS src/iTPS/transfer_matrix.{hpp,cpp}, src/arnoldi.cpp).
#pragma once

#include <complex>
#include <vector>

#include "ctm.hpp"

namespace tenes {
namespace itps {

// transfer_matrix.hpp:32-47 (defaults)
struct TransferMatrix_Parameters {
  bool to_calculate = true;
  int num_eigvals = 4;
  int maxdim_dense_eigensolver = 200;
  int arnoldi_maxdim = 50;
  int arnoldi_restartdim = 20;
  int arnoldi_maxiter = 1;
  double arnoldi_rtol = 1.0e-10;
};

class TransferMatrix_ctm {
 public:
  TransferMatrix_ctm(const SquareLattice &lattice,
                     const std::vector<tensor> &Tn,
                     const std::vector<tensor> &C1,
                     const std::vector<tensor> &C2,
                     const std::vector<tensor> &C3,
                     const std::vector<tensor> &C4,
                     const std::vector<tensor> &eTl,
                     const std::vector<tensor> &eTt,
                     const std::vector<tensor> &eTr,
                     const std::vector<tensor> &eTb)
      : lattice(lattice), Tn(Tn), C1(C1), C2(C2), C3(C3), C4(C4), eTl(eTl),
        eTt(eTt), eTr(eTr), eTb(eTb) {}

  // transfer_matrix.cpp:84-142 with the Arnoldi passes and restarts of
  // arnoldi.cpp:56-167 (maxiter = 1, the reference default, never restarts).
  // num_eigvals above arnoldi_maxdim on the Arnoldi path raises an error with
  // status -3 (the reference would index past its Krylov basis).
  std::vector<std::complex<double>> eigenvalues(
      int dir, int fixed_coord, const TransferMatrix_Parameters &params) const;

  size_t dim(int dir, int fixed_coord) const;
  tensor initial_vector(int dir, int fixed_coord) const;
  tensor matvec(int dir, const tensor &invec, int fixed_coord) const;
  tensor matrix(int dir, int fixed_coord) const;

 private:
  SquareLattice lattice;
  const std::vector<tensor> &Tn, &C1, &C2, &C3, &C4, &eTl, &eTt, &eTr, &eTb;
};

}  // namespace itps
}  // namespace tenes
