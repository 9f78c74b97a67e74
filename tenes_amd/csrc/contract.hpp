// CTM-environment contractions for observables and the full-update bond
// environment, with the call surface of TeNeS
// (src/iTPS/core/contract_itps_ctm.hpp:45-160, core/full_update.hpp:33-79).
#pragma once

#include <vector>

#include "tensor.hpp"

namespace tenes {
namespace itps {
namespace core {

using tensor = tn::Tensor;

double Contract_one_site_iTPS_CTM(const tensor &C1, const tensor &C2,
                                  const tensor &C3, const tensor &C4,
                                  const tensor &eT1, const tensor &eT2,
                                  const tensor &eT3, const tensor &eT4,
                                  const tensor &Tn1, const tensor &op1);

double Contract_two_sites_horizontal_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op1, const tensor &op2);

double Contract_two_sites_vertical_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op1, const tensor &op2);

double Contract_two_sites_horizontal_op12_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op12);

double Contract_two_sites_vertical_op12_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op12);

// Contract_iTPS_CTM dispatcher (contract_itps_ctm/ctm.cpp:45-87): every
// cluster of 1..4 rows and 1..4 columns.  1x1, 1x2, 2x1 and 2x2 follow the
// reference's trees; the other twelve shapes (ctm_1x3.cpp ... ctm_4x4.cpp)
// use one boundary sweep (contract_cluster_sweep).  Larger clusters raise an
// error, as the reference's dispatcher does.
double Contract_iTPS_CTM(const std::vector<const tensor *> &C,
                         const std::vector<const tensor *> &eTt,
                         const std::vector<const tensor *> &eTr,
                         const std::vector<const tensor *> &eTb,
                         const std::vector<const tensor *> &eTl,
                         const std::vector<std::vector<const tensor *>> &Tn,
                         const std::vector<std::vector<const tensor *>> &op);

// Any nrow x ncol cluster by a boundary sweep: the left column (C1, eTl...,
// C4) is carried across the columns, each absorbed as eTt, then ket and bra
// site by site, then eTb, and closed by the right column (or the same with
// rows and columns exchanged when nrow > ncol).  Peak intermediate:
// chi^2 (D^2)^min(nrow,ncol) * D^2 d entries.
double contract_cluster_sweep(const std::vector<const tensor *> &C,
                              const std::vector<const tensor *> &eTt,
                              const std::vector<const tensor *> &eTr,
                              const std::vector<const tensor *> &eTb,
                              const std::vector<const tensor *> &eTl,
                              const std::vector<std::vector<const tensor *>> &Tn,
                              const std::vector<std::vector<const tensor *>> &op);

tensor Create_Environment_two_sites(const tensor &C1, const tensor &C2,
                                    const tensor &C3, const tensor &C4,
                                    const tensor &eT1, const tensor &eT2,
                                    const tensor &eT3, const tensor &eT4,
                                    const tensor &eT5, const tensor &eT6,
                                    const tensor &Q1, const tensor &Q2);

// contract_itps_ctm/correlation.cpp:34-105: a row of the environment
// absorbed column by column, for <O(0) O(r)> along the row.
// A: (e1r, e3r, n1r, n2r).
void StartCorrelation_iTPS_CTM(tensor &A, const tensor &C1, const tensor &C4,
                               const tensor &eT1, const tensor &eT3,
                               const tensor &eT4, const tensor &Tn1,
                               const tensor &op);
void Transfer_iTPS_CTM(tensor &A, const tensor &eT1, const tensor &eT3,
                       const tensor &Tn1);
double FinishCorrelation_iTPS_CTM(const tensor &A, const tensor &C2,
                                  const tensor &C3, const tensor &eT1,
                                  const tensor &eT2, const tensor &eT3,
                                  const tensor &Tn1, const tensor &op);

}  // namespace core
}  // namespace itps
}  // namespace tenes
