// CTMRG core on the GPU, with the call surface of TeNeS's
// tenes::itps::core (src/iTPS/core/ctm.hpp:29-185): the same function names,
// argument order and in-place semantics, over tn::Tensor instead of
// mptensor::Tensor.  Only real (FP64) tensors are supported.
#pragma once

#include <array>
#include <vector>

#include "linalg.hpp"
#include "tensor.hpp"

namespace tenes {
namespace itps {

using tensor = tn::Tensor;

// Restatement of the CTM-relevant fields of PEPS_Parameters
// (src/iTPS/PEPS_Parameters.hpp:29-94, defaults PEPS_Parameters.cpp:29-77),
// plus the GPU SVD knobs.
struct PEPS_Parameters {
  int CHI = 2;
  double Inverse_projector_cut = 1e-12;
  double CTM_Convergence_Epsilon = 1e-6;
  int Max_CTM_Iteration = 100;
  bool CTM_Projector_corner = true;
  bool Use_RSVD = false;
  double RSVD_Oversampling_factor = 2.0;
  int print_level = 1;
  // GPU truncated-SVD controls (no counterpart in the reference).
  // Range-finder iterations from a Gaussian start block (full-SVD mode).
  int svd_power_iters = 8;
  int rsvd_power_iters = 1;
  // Warm start: the range finder of each projector starts from the right
  // Ritz basis the same projector (environment, move, coordinate, row)
  // produced in its previous sweep, kept across Calc_CTM_Environment calls
  // that continue an environment (initialize = false), and runs
  // svd_warm_power_iters iterations.  Heisenberg D=5, chi=100, 12 sweeps:
  // 8 iterations, warm or cold, put the correlation lengths at the level
  // where any converged run lands (2.3e-8 from the exact-SVD run), warm at
  // 90.6 instead of 96.3 ms per sweep; 4 iterations leave 3e-7 to 1.2e-6
  // (seed dependent) at 72-76 ms per sweep.
  bool svd_warm_start = true;
  int svd_warm_power_iters = 8;
  // Mixed precision, the contractions of the four moves on the FP32 MFMA:
  // 1 = enlarged corners, projectors, Calc_Next_CTM / Calc_Next_eT, with the
  // matrix the SVD truncates (R = LT LB) kept FP64; 2 = R as well; 3 = all
  // of them, R included, on the bf16 MFMA as three products of hi/lo bf16
  // operand halves with FP32 accumulation.  The projector SVD and its QRs,
  // the convergence check and everything outside the moves stay FP64.
  // 0 = all FP64.
  int contraction_fp32 = 0;
  // Where the projector cut of ctm.cpp:275-288 is decided.  On the host
  // (default): the reference's variable shapes, one synchronisation per
  // projector.  On the device (no counterpart in the reference): every
  // projector keeps the full min(rows, cols, CHI) columns and the dropped
  // directions (s_i / s_0 <= Inverse_projector_cut) get a zero column, so a
  // sweep needs no host round trip and its shapes are fixed; the environment
  // tensors then stay CHI-sized where the reference's shrink (zero rows and
  // columns in their place: the same network).  Measured at D=5, chi=100
  // with the overlapped move chains: 16.9 sweeps/s on the device against
  // 20.5 on the host, with the same kernels; the fixed shapes are what a
  // captured sweep needs.
  bool projector_cut_on_device = false;
  // Sweeps of the environment driver with corner projectors captured into a
  // hipGraph (no counterpart in the reference): the cut goes on the device,
  // the first sweep with every projector warm-started is captured, later
  // sweeps replay the graph while the environment's buffers, shapes and
  // warm-start decisions stay the same (else it is captured again), and a
  // sweep the graph cannot cover runs as usual.
  bool graph_sweeps = false;
};

// Identity of an environment for the warm-start bases of its projectors:
// the address of its C1 array (stable while the environment lives).
inline int64_t environment_id(const std::vector<tn::Tensor> &C1) {
  return (int64_t)(intptr_t)&C1;
}

// Restatement of SquareLattice (src/SquareLattice.{hpp,cpp}): site index
// x + y*LX, neighbours 0:left 1:top 2:right 3:bottom, skew boundary.
class SquareLattice {
 public:
  int LX = 1, LX_noskew = 1, LY = 1, LY_noskew = 1, N_UNIT = 1,
      N_UNIT_noskew = 1, skew = 0;
  SquareLattice() = default;
  SquareLattice(int X, int Y, int skew = 0);
  int x(int i) const { return i % LX; }
  int y(int i) const { return i / LX; }
  int index(int x, int y) const;
  int neighbor(int i, int dir) const { return nn_[i][dir]; }
  int left(int i) const { return nn_[i][0]; }
  int top(int i) const { return nn_[i][1]; }
  int right(int i) const { return nn_[i][2]; }
  int bottom(int i) const { return nn_[i][3]; }
  int other(int index, int dx, int dy) const;
  int parity(int i) const { return parity_[i]; }

 private:
  std::vector<std::array<int, 4>> nn_;
  std::vector<int> parity_;
};

namespace core {

void Calc_projector_left_block(const tensor &C1, const tensor &C4,
                               const tensor &eT1, const tensor &eT6,
                               const tensor &eT7, const tensor &eT8,
                               const tensor &Tn1, const tensor &Tn4,
                               const PEPS_Parameters &peps_parameters,
                               tensor &PU, tensor &PL);

void Calc_projector_updown_blocks(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &eT7, const tensor &eT8,
    const tensor &Tn1, const tensor &Tn2, const tensor &Tn3, const tensor &Tn4,
    const PEPS_Parameters &peps_parameters, tensor &PU, tensor &PL);

void Calc_Next_CTM(const tensor &C1, const tensor &C4, const tensor &eT1,
                   const tensor &eT6, const tensor &PU, const tensor &PL,
                   tensor &C1_out, tensor &C4_out);

void Calc_Next_eT(const tensor &eT8, const tensor &Tn1, const tensor &PU,
                  const tensor &PL, tensor &eT_out);

void Left_move(std::vector<tensor> &C1, const std::vector<tensor> &C2,
               const std::vector<tensor> &C3, std::vector<tensor> &C4,
               const std::vector<tensor> &eTt, const std::vector<tensor> &eTr,
               const std::vector<tensor> &eTb, std::vector<tensor> &eTl,
               const std::vector<tensor> &Tn, int ix,
               const PEPS_Parameters &peps_parameters,
               const SquareLattice &lattice);
void Right_move(const std::vector<tensor> &C1, std::vector<tensor> &C2,
                std::vector<tensor> &C3, const std::vector<tensor> &C4,
                const std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                const std::vector<tensor> &eTb, const std::vector<tensor> &eTl,
                const std::vector<tensor> &Tn, int ix,
                const PEPS_Parameters &peps_parameters,
                const SquareLattice &lattice);
void Top_move(std::vector<tensor> &C1, std::vector<tensor> &C2,
              const std::vector<tensor> &C3, const std::vector<tensor> &C4,
              std::vector<tensor> &eTt, const std::vector<tensor> &eTr,
              const std::vector<tensor> &eTb, const std::vector<tensor> &eTl,
              const std::vector<tensor> &Tn, int iy,
              const PEPS_Parameters &peps_parameters,
              const SquareLattice &lattice);
void Bottom_move(const std::vector<tensor> &C1, const std::vector<tensor> &C2,
                 std::vector<tensor> &C3, std::vector<tensor> &C4,
                 const std::vector<tensor> &eTt,
                 const std::vector<tensor> &eTr, std::vector<tensor> &eTb,
                 const std::vector<tensor> &eTl, const std::vector<tensor> &Tn,
                 int iy, const PEPS_Parameters &peps_parameters,
                 const SquareLattice &lattice);

// Convergence test of ctm.cpp:815-969 (same early-exit order).  The
// L2-normalised singular values of C_old are taken from `old_spectra` when
// provided (they equal the previous sweep's new spectra), and the new
// spectra are returned through `new_spectra`.
bool Check_Convergence_CTM(const std::vector<tensor> &C1,
                           const std::vector<tensor> &C2,
                           const std::vector<tensor> &C3,
                           const std::vector<tensor> &C4,
                           const std::vector<tensor> &C1_old,
                           const std::vector<tensor> &C2_old,
                           const std::vector<tensor> &C3_old,
                           const std::vector<tensor> &C4_old,
                           const PEPS_Parameters &peps_parameters,
                           const SquareLattice &lattice, double &sig_max,
                           double *sig_all = nullptr);

// Initial C/eT from the site tensors (ctm.cpp:984-1193).
void Initialize_CTM(std::vector<tensor> &C1, std::vector<tensor> &C2,
                    std::vector<tensor> &C3, std::vector<tensor> &C4,
                    std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                    std::vector<tensor> &eTb, std::vector<tensor> &eTl,
                    const std::vector<tensor> &Tn,
                    const PEPS_Parameters &peps_parameters,
                    const SquareLattice &lattice);

// ctm.cpp:971-1253.  Returns the number of sweeps; *sig_max_out (if given)
// receives the last convergence measure.
int Calc_CTM_Environment(std::vector<tensor> &C1, std::vector<tensor> &C2,
                         std::vector<tensor> &C3, std::vector<tensor> &C4,
                         std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                         std::vector<tensor> &eTb, std::vector<tensor> &eTl,
                         const std::vector<tensor> &Tn,
                         const PEPS_Parameters &peps_parameters,
                         const SquareLattice &lattice, bool initialize = true,
                         double *sig_max_out = nullptr);

// The arrays of one environment: C1..C4, eTt, eTr, eTb, eTl, and its site
// tensors and lattice.
struct CtmEnvironment {
  std::vector<tensor> *A[8];
  const std::vector<tensor> *Tn;
  const SquareLattice *lat;
};

// Calc_CTM_Environment for several environments of one context at once:
// every environment runs the reference's own sequence of sweeps and
// convergence checks (with its own parameters), and the environments
// advance in lockstep, one sweep each per round, so that one batched
// convergence check and one synchronisation serve the round and captured
// sweeps (graph_sweeps) of different environments replay side by side.
// Returns each environment's sweep count.
std::vector<int> Calc_CTM_Environment_lockstep(std::vector<CtmEnvironment> &envs,
                                               const std::vector<PEPS_Parameters> &pps,
                                               bool initialize,
                                               std::vector<double> *sig_max = nullptr);

// Singular spectra of C (L2 normalised, descending) for diagnostics/tests.
std::vector<double> normalized_spectrum(const tensor &C);

}  // namespace core
}  // namespace itps
}  // namespace tenes
