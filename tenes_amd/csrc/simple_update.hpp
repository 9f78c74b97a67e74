// Simple update on the GPU: produces the site tensors the CTM path consumes
// (TeNeS src/iTPS/core/simple_update.cpp:50-198 and the driver loop of
// src/iTPS/simple_update.cpp:103-138).
#pragma once

#include <vector>

#include "ctm.hpp"

namespace tenes {
namespace itps {
namespace core {

// lambda1/lambda2: the four bond-weight vectors of each site (leg order).
void Simple_update_bond(const tensor &Tn1, const tensor &Tn2,
                        const std::vector<std::vector<double>> &lambda1,
                        const std::vector<std::vector<double>> &lambda2,
                        const tensor &op12, int connect1,
                        double inverse_lambda_cut, tensor &Tn1_new,
                        tensor &Tn2_new, std::vector<double> &lambda_c);

struct EvolutionOp {
  int source_site;
  int source_leg;
  tensor op;  // (d1, d2, d1, d2)
};

// nsteps passes over `ops` (group 0, no local gauge fixing).
void Simple_update_steps(std::vector<tensor> &Tn,
                         std::vector<std::vector<std::vector<double>>> &lambda,
                         const SquareLattice &lattice,
                         const std::vector<EvolutionOp> &ops, int nsteps,
                         double inverse_lambda_cut);

}  // namespace core
}  // namespace itps
}  // namespace tenes
