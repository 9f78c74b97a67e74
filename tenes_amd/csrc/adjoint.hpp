// Bond reduced density matrices with the CTM environment held fixed, and
// their adjoint with respect to the site tensors (dE/dTn for an energy made
// of nearest-neighbour bond terms).
//
// The forward pass is the contraction tree of
// Contract_two_sites_{horizontal,vertical}_op12_iTPS_CTM
// (src/iTPS/core/contract_itps_ctm/ctm.cpp:213-319) stopped before the
// operator: the two halves of the bond (one per site) contracted over their
// shared legs give rho[p, q, t, u] = (ket1, bra1, ket2, bra2), and
// <op12> = sum op12[p, t, q, u] rho[p, q, t, u].  The reverse pass walks the
// same tree backwards (each tensordot's adjoint is two more tensordots on the
// GEMM kernels); a BondTape keeps the intermediates of one bond only, so a
// sum over bonds recomputes per bond instead of storing every bond's tape.
// The gradient has no counterpart in the reference (TeNeS optimises by
// imaginary-time evolution).
#pragma once

#include <vector>

#include "tensor.hpp"

namespace tenes {
namespace itps {
namespace core {

using tensor = tn::Tensor;

struct BondHalf {
  const tensor *T = nullptr;  // site tensor (ket; the bra is conj(T) = T)
  tensor e;                   // edge tensor contracted last
  tensor a3, a4, a5, full;    // before the bra, after the bra, after the ket
  tn::Axes ax4A, ax4B, ax5A, ax5B, ax6A, ax6B;
};

struct BondTape {
  BondHalf h1, h2;
  tensor rho;  // (d1, d1, d2, d2) = (ket1, bra1, ket2, bra2)
};

// Horizontal bond Tn1 (left) - Tn2 (right); arguments as
// Contract_two_sites_horizontal_op12_iTPS_CTM (contract_itps_ctm/ctm.cpp:213).
BondTape bond_rdm_horizontal(const tensor &C1, const tensor &C2,
                             const tensor &C3, const tensor &C4,
                             const tensor &eT1, const tensor &eT2,
                             const tensor &eT3, const tensor &eT4,
                             const tensor &eT5, const tensor &eT6,
                             const tensor &Tn1, const tensor &Tn2);

// Vertical bond Tn1 (top) - Tn2 (bottom); arguments as
// Contract_two_sites_vertical_op12_iTPS_CTM (contract_itps_ctm/ctm.cpp:268).
BondTape bond_rdm_vertical(const tensor &C1, const tensor &C2,
                           const tensor &C3, const tensor &C4,
                           const tensor &eT1, const tensor &eT2,
                           const tensor &eT3, const tensor &eT4,
                           const tensor &eT5, const tensor &eT6,
                           const tensor &Tn1, const tensor &Tn2);

// G1 += d<rho, dRho>/dTn1, G2 += d<rho, dRho>/dTn2 (ket and bra occurrences;
// G1 and G2 may be the same tensor when both sites are one tensor).
void bond_rdm_adjoint(const BondTape &tape, const tensor &dRho, tensor &G1,
                      tensor &G2);

// One nearest-neighbour bond term: sites s1 (left / top) and s2 (right /
// bottom) and its operator in (s1, s2, s1', s2') order, the op12 argument of
// Contract_two_sites_*_op12_iTPS_CTM.
struct BondTerm {
  int s1 = 0, s2 = 0;
  bool vertical = false;
  std::vector<double> op;  // d1 * d2 * d1 * d2, row-major
};

// E = (1/n_sites) sum_b <op_b>, <op_b> = sum op[p,t,q,u] rho[p,q,t,u] /
// sum rho[p,p,t,t] (the bond values of twosite_obs.cpp:146-262 summed as
// density.cpp:39-76 does), with env = {C1, C2, C3, C4, eTt, eTr, eTb, eTl,
// Tn} arrays indexed by site.  If G is non-null, (*G)[i] (shaped like Tn[i])
// is overwritten with dE/dTn_i: per bond, the adjoint of rho with
// dRho[p,q,t,u] = (op[p,t,q,u] - <op_b> delta_pq delta_tu) / norm_b / n_sites.
double bond_energy(const std::vector<tensor> *env,
                   const std::vector<BondTerm> &terms, int n_sites,
                   std::vector<tensor> *G);

}  // namespace core
}  // namespace itps
}  // namespace tenes
