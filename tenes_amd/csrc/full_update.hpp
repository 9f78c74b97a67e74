// Full update of one bond with the CTM environment (TeNeS
// src/iTPS/core/full_update.{hpp,cpp}: Full_update_bond_horizontal 85-393,
// Full_update_bond 395-441) and the fast-full-update driver step
// (src/iTPS/full_update.cpp:29-186).
//
// The bond environment N(r1, r2, r1*, r2*) (Create_Environment_two_sites,
// core/full_update.cpp:51-83: chi^3 D^4-sized contractions) runs on the GPU;
// everything after it works on tensors of at most (D d)^4 entries and runs in
// FP64 on the host (small_dense.hpp), where a kernel launch would cost more
// than the arithmetic.  The fast-full-update re-moves are the GPU CTM moves.
#pragma once

#include <vector>

#include "ctm.hpp"
#include "small_dense.hpp"

namespace tenes {
namespace itps {

// The full-update fields of PEPS_Parameters (PEPS_Parameters.hpp:29-94,
// defaults PEPS_Parameters.cpp:55-64).
struct FU_Parameters {
  double Inverse_Env_cut = 1e-12;
  double Full_Inverse_precision = 1e-12;
  double Full_Convergence_Epsilon = 1e-6;
  int Full_max_iteration = 100;
  bool Full_Gauge_Fix = true;
  bool Full_Use_FastFullUpdate = true;
};

namespace core {

// core/full_update.cpp:85-393: op12 is a host tensor (d1, d2, d1o, d2o) with
// d1, d2 the physical dimensions of Tn1, Tn2; the new tensors get d1o, d2o
// (a path-decomposed gate changes the intermediate site's physical leg).
void Full_update_bond_horizontal(const tensor &C1, const tensor &C2,
                                 const tensor &C3, const tensor &C4,
                                 const tensor &eT1, const tensor &eT2,
                                 const tensor &eT3, const tensor &eT4,
                                 const tensor &eT5, const tensor &eT6,
                                 const tensor &Tn1, const tensor &Tn2,
                                 const tn::host::HT &op12,
                                 const FU_Parameters &fp, tensor &Tn1_new,
                                 tensor &Tn2_new);

// core/full_update.cpp:395-441: connect1 = leg of Tn1 on the bond.
void Full_update_bond(const tensor &C1, const tensor &C2, const tensor &C3,
                      const tensor &C4, const tensor &eT1, const tensor &eT2,
                      const tensor &eT3, const tensor &eT4, const tensor &eT5,
                      const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
                      const tn::host::HT &op12, int connect1,
                      const FU_Parameters &fp, tensor &Tn1_new,
                      tensor &Tn2_new);

}  // namespace core

// src/iTPS/full_update.cpp:29-152 for one two-site evolution operator on the
// bond (source, source_leg): the bond update, then the fast-full-update
// re-moves of the two columns / rows (or a fresh CTM environment).
void full_update_step(std::vector<tensor> &C1, std::vector<tensor> &C2,
                      std::vector<tensor> &C3, std::vector<tensor> &C4,
                      std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                      std::vector<tensor> &eTb, std::vector<tensor> &eTl,
                      std::vector<tensor> &Tn, int source, int source_leg,
                      const tn::host::HT &op12,
                      const PEPS_Parameters &pp, const FU_Parameters &fp,
                      const SquareLattice &lat);

}  // namespace itps
}  // namespace tenes
