// Checkpoint of an environment: site tensors, CTM environment and the
// simple-update bond weights, in the directory layout of TeNeS's
// iTPS::save_tensors / load_tensors (src/iTPS/saveload_tensors.cpp:38-230).
#pragma once

#include <string>
#include <vector>

#include "tensor.hpp"

namespace tenes {
namespace itps {

// The nine tensor kinds of one site in the environment's order (C1..C4,
// eTt, eTr, eTb, eTl, Tn) and their file stems (saveload_tensors.cpp:61-70).
constexpr int kCheckpointKinds = 9;

struct CheckpointState {
  std::vector<tn::Tensor> *arr;  // [kCheckpointKinds], each N_UNIT long
  std::vector<std::vector<std::vector<double>>> *lambda;  // [site][leg]
  int n_unit;
};

// Writes <dir>/params.dat (format version 1: N_UNIT, CHI and the shape of
// every site tensor), <dir>/<stem>_<i>.dat for every tensor and
// <dir>/lambda_<i>.dat (the four legs' weights, one per line).  The
// directory is created when it does not exist.  Every tensor must be set.
void save_tensors(const CheckpointState &st, const std::string &dir, int chi);

// Reads a checkpoint written by save_tensors into st, each tensor resized to
// the shape the environment already holds for it (zero padding of shorter
// legs, then cutting of longer ones, as resize_tensor, src/tensor.cpp:
// 114-148); a kind the environment does not hold yet keeps the loaded
// shape, except that chi > 0 resizes the two environment legs of C and eT
// to chi.  The bond weights are padded with zeros or cut to the site
// tensors' bond dimensions.  Errors (as saveload_tensors.cpp:88-230): the
// directory does not exist, no params.dat (format version 0, not
// supported), an unknown format version, another N_UNIT, another physical
// dimension, a tensor of another rank.
void load_tensors(CheckpointState &st, tn::Context &ctx, const std::string &dir,
                  int chi);

}  // namespace itps
}  // namespace tenes
