// extern "C" boundary: handles around tn::Context / tn::Tensor and the
// environment arrays, exception-free, one status code per call.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tenes_amd.h"
#include "adjoint.hpp"
#include "checkpoint.hpp"
#include "contract.hpp"
#include "ctm.hpp"
#include "full_update.hpp"
#include "linalg.hpp"
#include "noalloc.hpp"
#include "simple_update.hpp"
#include "small_dense.hpp"
#include "transfer.hpp"

struct tenes_ctx {
  std::unique_ptr<tn::Context> c;
  std::string err;
};

struct tenes_tensor {
  tn::Tensor t;
  tenes_ctx *owner;
};

struct tenes_env {
  tenes_ctx *ctx;
  tenes::itps::SquareLattice lat;
  std::vector<tn::Tensor> arr[9];  // C1..C4, eTt, eTr, eTb, eTl, Tn
  std::vector<tn::Tensor> saved[4];
  std::vector<std::vector<std::vector<double>>> lambda;  // [site][leg]
};

namespace {

using tenes::itps::PEPS_Parameters;

thread_local std::string g_noctx_err;

template <class F>
int guard(tenes_ctx *ctx, F &&f) {
  try {
    f();
    return TENES_OK;
  } catch (const tn::Error &e) {
    (ctx ? ctx->err : g_noctx_err) = e.what();
    return e.code < 0 ? e.code : TENES_ERR_GENERIC;
  } catch (const std::exception &e) {
    (ctx ? ctx->err : g_noctx_err) = e.what();
    return TENES_ERR_GENERIC;
  } catch (...) {
    (ctx ? ctx->err : g_noctx_err) = "unknown error";
    return TENES_ERR_GENERIC;
  }
}

tenes_tensor *wrap(tenes_ctx *ctx, const tn::Tensor &t) {
  return new tenes_tensor{t, ctx};
}

tn::Shape to_shape(int rank, const int64_t *shape) {
  TN_REQUIRE(rank >= 0 && rank <= 16, "rank out of range");
  TN_REQUIRE(rank == 0 || shape != nullptr, "null shape");
  return tn::Shape(shape, shape + rank);
}

PEPS_Parameters to_params(const tenes_peps_parameters *p) {
  PEPS_Parameters q;
  if (!p) return q;
  q.CHI = p->CHI;
  q.Inverse_projector_cut = p->Inverse_projector_cut;
  q.CTM_Convergence_Epsilon = p->CTM_Convergence_Epsilon;
  q.Max_CTM_Iteration = p->Max_CTM_Iteration;
  q.CTM_Projector_corner = p->CTM_Projector_corner != 0;
  q.Use_RSVD = p->Use_RSVD != 0;
  q.RSVD_Oversampling_factor = p->RSVD_Oversampling_factor;
  q.print_level = p->print_level;
  q.svd_power_iters = p->svd_power_iters;
  q.rsvd_power_iters = p->rsvd_power_iters;
  q.svd_warm_start = p->svd_warm_start != 0;
  q.svd_warm_power_iters = p->svd_warm_power_iters;
  q.contraction_fp32 = p->contraction_fp32;
  TN_REQUIRE(q.contraction_fp32 >= 0 && q.contraction_fp32 <= 3,
             "contraction_fp32 must be 0, 1, 2 or 3");
  q.projector_cut_on_device = p->projector_cut_on_device != 0;
  q.graph_sweeps = p->graph_sweeps != 0;
  TN_REQUIRE(q.CHI >= 1, "CHI must be positive");
  TN_REQUIRE(q.RSVD_Oversampling_factor >= 1.0,
             "RSVD oversampling factor must be >= 1");
  return q;
}

const tn::Tensor &T(const tenes_tensor *t) {
  TN_REQUIRE(t != nullptr && t->t.defined(), "null tensor handle");
  return t->t;
}

tenes_ctx *ctx_of(const tenes_tensor *t) {
  TN_REQUIRE(t != nullptr, "null tensor handle");
  return t->owner;
}

tn::Tensor host_op(tenes_ctx *ctx, const double *op, int64_t d) {
  TN_REQUIRE(op != nullptr, "null operator");
  return tn::from_host(*ctx->c, tn::Shape{d, d}, op);
}

tn::Tensor identity_op(tenes_ctx *ctx, int64_t d) {
  std::vector<double> id((size_t)(d * d), 0.0);
  for (int64_t i = 0; i < d; ++i) id[(size_t)(i * d + i)] = 1.0;
  return tn::from_host(*ctx->c, tn::Shape{d, d}, id.data());
}

// iTPO environments hold site tensors (D,D,D,D,d,d) (phys_ket, phys_bra)
// and are contracted as TeNeS's single-layer networks
// (contract_density_ctm/, ctm_single.cpp): every site enters as its TPO
// traced against the one-site operator placed on it, sum_ab T[..,a,b]
// op[a,b], shaped (D,D,D,D,1), under the context's single-layer mode and
// with a 1x1 identity in the operator's place.
bool is_tpo(const tenes_env *env) {
  return env->arr[TENES_TN][0].rank() == 6;
}

tn::Tensor tpo_ket(const tn::Tensor &T, const tn::Tensor &op) {
  TN_REQUIRE(T.rank() == 6 && op.rank() == 2 && op.shape[0] == T.shape[4] &&
                 op.shape[1] == T.shape[5],
             "TPO site tensor and operator do not match");
  return tn::reshape(tn::tensordot(T, op, tn::Axes{4, 5}, tn::Axes{0, 1}),
                     tn::Shape{T.shape[0], T.shape[1], T.shape[2], T.shape[3], 1});
}

using Grid = std::vector<std::vector<const tn::Tensor *>>;

// Contract_iTPS_CTM, or Contract_density_CTM (contract_density_ctm/ctm.cpp:
// 45-85) for an iTPO environment.
double cluster_value(tenes_env *env, const std::vector<const tn::Tensor *> &C,
                     const std::vector<const tn::Tensor *> &eTt,
                     const std::vector<const tn::Tensor *> &eTr,
                     const std::vector<const tn::Tensor *> &eTb,
                     const std::vector<const tn::Tensor *> &eTl,
                     const Grid &Tn, const Grid &op) {
  namespace K = tenes::itps::core;
  if (!is_tpo(env)) return K::Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, Tn, op);
  std::deque<tn::Tensor> keep;
  keep.push_back(identity_op(env->ctx, 1));
  const tn::Tensor *one = &keep.back();
  Grid kets = Tn, ones = op;
  for (size_t r = 0; r < Tn.size(); ++r)
    for (size_t c = 0; c < Tn[r].size(); ++c) {
      keep.push_back(tpo_ket(*Tn[r][c], *op[r][c]));
      kets[r][c] = &keep.back();
      ones[r][c] = one;
    }
  tn::SingleLayer single(*env->ctx->c, true);
  return K::Contract_iTPS_CTM(C, eTt, eTr, eTb, eTl, kets, ones);
}

void check_env(const tenes_env *env) {
  TN_REQUIRE(env != nullptr, "null environment");
  for (int k = 0; k < 9; ++k)
    for (const auto &t : env->arr[k])
      TN_REQUIRE(t.defined(), "environment tensor not set (kind " +
                                  std::to_string(k) + ")");
}

// indices[row][col] of twosite_obs.cpp:119-176 for a bond (source, dx, dy)
struct Cluster {
  int nrow, ncol, source_row, source_col, target_row, target_col;
  std::vector<std::vector<int>> idx;
};

Cluster cluster_of(const tenes::itps::SquareLattice &lat, int source, int dx,
                   int dy) {
  Cluster c;
  c.ncol = std::abs(dx) + 1;
  c.nrow = std::abs(dy) + 1;
  if (dx >= 0) {
    c.source_col = 0;
    c.target_col = c.ncol - 1;
  } else {
    c.source_col = c.ncol - 1;
    c.target_col = 0;
  }
  if (dy >= 0) {
    c.source_row = c.nrow - 1;
    c.target_row = 0;
  } else {
    c.source_row = 0;
    c.target_row = c.nrow - 1;
  }
  c.idx.assign(c.nrow, std::vector<int>(c.ncol));
  for (int row = 0; row < c.nrow; ++row)
    for (int col = 0; col < c.ncol; ++col)
      c.idx[row][col] =
          lat.other(source, col - c.source_col, c.source_row - row);
  return c;
}


// The cluster of a bond (twosite_obs.cpp:119-260) with one-site operators on
// its source and target (nullptr = identity).
double bond_cluster_value(tenes_env *env, const Cluster &c,
                          const tn::Tensor *op_source,
                          const tn::Tensor *op_target) {
  auto &a = env->arr;
  std::vector<const tn::Tensor *> C(4), eTt(c.ncol), eTr(c.nrow), eTb(c.ncol),
      eTl(c.nrow);
  Grid Tn(c.nrow, std::vector<const tn::Tensor *>(c.ncol)),
      op(c.nrow, std::vector<const tn::Tensor *>(c.ncol));
  std::deque<tn::Tensor> keep;
  for (int row = 0; row < c.nrow; ++row) {
    for (int col = 0; col < c.ncol; ++col) {
      const int idx = c.idx[row][col];
      Tn[row][col] = &a[8][idx];
      keep.push_back(identity_op(env->ctx, a[8][idx].shape[4]));
      op[row][col] = &keep.back();
    }
    eTl[row] = &a[7][c.idx[row][0]];
    eTr[row] = &a[5][c.idx[row][c.ncol - 1]];
  }
  for (int col = 0; col < c.ncol; ++col) {
    eTt[col] = &a[4][c.idx[0][col]];
    eTb[col] = &a[6][c.idx[c.nrow - 1][col]];
  }
  C[0] = &a[0][c.idx[0][0]];
  C[1] = &a[1][c.idx[0][c.ncol - 1]];
  C[2] = &a[2][c.idx[c.nrow - 1][c.ncol - 1]];
  C[3] = &a[3][c.idx[c.nrow - 1][0]];
  if (op_source) op[c.source_row][c.source_col] = op_source;
  if (op_target) op[c.target_row][c.target_col] = op_target;
  return cluster_value(env, C, eTt, eTr, eTb, eTl, Tn, op);
}

}  // namespace

extern "C" {

int tenes_abi_version(void) { return 207; }

int tenes_ctx_create(int device, void *hip_stream, tenes_ctx **out) {
  return guard(nullptr, [&] {
    TN_REQUIRE(out != nullptr, "null output pointer");
    auto *c = new tenes_ctx;
    try {
      c->c.reset(new tn::Context(device, static_cast<hipStream_t>(hip_stream)));
      // the allocation-free entry point's scratch address on this device
      // (the context constructor made the device current)
      tn::noalloc::init_device(device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int tenes_ctx_destroy(tenes_ctx *ctx) {
  delete ctx;
  return TENES_OK;
}

const char *tenes_ctx_last_error(const tenes_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_noctx_err.c_str();
}

int tenes_ctx_synchronize(tenes_ctx *ctx) {
  return guard(ctx, [&] { ctx->c->sync(); });
}

int tenes_ctx_set_seed(tenes_ctx *ctx, uint64_t seed) {
  return guard(ctx, [&] { ctx->c->seed = seed; });
}

int tenes_ctx_stats(tenes_ctx *ctx, double *gemm_flops, int64_t *gemm_calls) {
  return guard(ctx, [&] {
    if (gemm_flops) *gemm_flops = ctx->c->gemm_flops;
    if (gemm_calls) *gemm_calls = ctx->c->gemm_calls;
  });
}

int tenes_ctx_reset_stats(tenes_ctx *ctx) {
  return guard(ctx, [&] {
    ctx->c->gemm_flops = 0.0;
    ctx->c->gemm_calls = 0;
    ctx->c->svd_calls = 0;
    ctx->c->jacobi_rounds = 0;
  });
}

int tenes_ctx_set_fp32_contractions(tenes_ctx *ctx, int on) {
  if (!ctx) return TENES_ERR_ARGUMENT;
  return guard(ctx, [&] {
    TN_REQUIRE(on >= 0 && on <= 2, "contraction mode must be 0, 1 or 2");
    ctx->c->set_contraction_mode(on);
  });
}

int tenes_ctx_set_fused_site(tenes_ctx *ctx, int on) {
  if (!ctx) return TENES_ERR_ARGUMENT;
  return guard(ctx, [&] { ctx->c->set_fused_site(on != 0); });
}

int tenes_ctx_set_autotune(tenes_ctx *ctx, int on) {
  if (!ctx) return TENES_ERR_ARGUMENT;
  return guard(ctx, [&] { ctx->c->set_autotune(on != 0); });
}

int tenes_ctx_set_stream_k(tenes_ctx *ctx, int on) {
  if (!ctx) return TENES_ERR_ARGUMENT;
  return guard(ctx, [&] { ctx->c->set_stream_k(on != 0); });
}

int tenes_ctx_set_max_lanes(tenes_ctx *ctx, int n) {
  return guard(ctx, [&] {
    TN_REQUIRE(n >= 1, "max lanes must be >= 1");
    ctx->c->set_max_lanes(n);
  });
}

int tenes_ctx_graph_stats(tenes_ctx *ctx, int64_t *captures,
                          int64_t *replays) {
  return guard(ctx, [&] {
    if (captures) *captures = ctx->c->graph_captures;
    if (replays) *replays = ctx->c->graph_replays;
  });
}

int tenes_ctx_set_max_moves_in_flight(tenes_ctx *ctx, int n) {
  return guard(ctx, [&] {
    TN_REQUIRE(n >= 0, "moves in flight must be >= 0");
    ctx->c->set_max_moves_in_flight(n);
  });
}

int tenes_ctx_lane_state(tenes_ctx *ctx, int *forked, int *lane) {
  if (!ctx) return TENES_ERR_ARGUMENT;
  return guard(ctx, [&] {
    if (forked) *forked = ctx->c->forked() ? 1 : 0;
    if (lane) *lane = ctx->c->current_lane();
  });
}

int tenes_ctx_svd_stats(tenes_ctx *ctx, int64_t *svd_calls,
                        int64_t *jacobi_rounds) {
  return guard(ctx, [&] {
    if (svd_calls) *svd_calls = ctx->c->svd_calls;
    if (jacobi_rounds) *jacobi_rounds = ctx->c->jacobi_rounds;
  });
}

int tenes_ctx_set_profiling(tenes_ctx *ctx, int on) {
  return guard(ctx, [&] { ctx->c->set_profiling(on != 0); });
}

int tenes_ctx_kernel_report(tenes_ctx *ctx, char *buf, int buflen) {
  return guard(ctx, [&] {
    TN_REQUIRE(buf && buflen > 0, "null buffer");
    std::string out;
    char line[512];
    for (auto &kv : ctx->c->kernel_stats()) {
      std::snprintf(line, sizeof(line), "%s %lld %.9g %.9g\n", kv.first.c_str(),
                    (long long)kv.second.launches, kv.second.ms,
                    kv.second.flops);
      out += line;
    }
    TN_REQUIRE((int)out.size() < buflen, "report buffer too small");
    std::memcpy(buf, out.c_str(), out.size() + 1);
  });
}

int tenes_ctx_reset_kernel_report(tenes_ctx *ctx) {
  return guard(ctx, [&] { ctx->c->reset_kernel_stats(); });
}

int tenes_tensor_from_host(tenes_ctx *ctx, int rank, const int64_t *shape,
                           const double *data, tenes_tensor **out) {
  return guard(ctx, [&] {
    TN_REQUIRE(out && data, "null pointer");
    *out = wrap(ctx, tn::from_host(*ctx->c, to_shape(rank, shape), data));
  });
}

int tenes_tensor_from_device(tenes_ctx *ctx, int rank, const int64_t *shape,
                             double *dev_ptr, tenes_tensor **out) {
  return guard(ctx, [&] {
    TN_REQUIRE(out && dev_ptr, "null pointer");
    *out = wrap(ctx, tn::borrow(*ctx->c, to_shape(rank, shape), dev_ptr));
  });
}

int tenes_tensor_zeros(tenes_ctx *ctx, int rank, const int64_t *shape,
                       tenes_tensor **out) {
  return guard(ctx, [&] {
    TN_REQUIRE(out, "null pointer");
    *out = wrap(ctx, tn::zeros(*ctx->c, to_shape(rank, shape)));
  });
}

int tenes_tensor_rank(const tenes_tensor *t, int *rank) {
  if (!t || !rank) return TENES_ERR_ARGUMENT;
  *rank = t->t.rank();
  return TENES_OK;
}

int tenes_tensor_shape(const tenes_tensor *t, int64_t *shape) {
  if (!t || !shape) return TENES_ERR_ARGUMENT;
  for (int i = 0; i < t->t.rank(); ++i) shape[i] = t->t.shape[i];
  return TENES_OK;
}

int tenes_tensor_to_host(const tenes_tensor *t, double *out) {
  if (!t) return TENES_ERR_ARGUMENT;
  return guard(t->owner, [&] {
    TN_REQUIRE(out != nullptr, "null output");
    tn::to_host(t->t, out);
  });
}

int tenes_tensor_device_ptr(const tenes_tensor *t, double **out) {
  if (!t || !out) return TENES_ERR_ARGUMENT;
  *out = t->t.data();
  return TENES_OK;
}

int tenes_tensor_destroy(tenes_tensor *t) {
  delete t;
  return TENES_OK;
}

int tenes_tensordot(const tenes_tensor *A, const tenes_tensor *B, int nax,
                    const int *axA, const int *axB, const int *perm,
                    tenes_tensor **out) {
  if (!A || !B) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(A);
  return guard(ctx, [&] {
    TN_REQUIRE(out != nullptr, "null output");
    TN_REQUIRE(nax >= 0 && (nax == 0 || (axA && axB)), "bad axes");
    tn::Axes a(axA, axA + nax), b(axB, axB + nax), pm;
    if (perm) {
      const int rr = T(A).rank() + T(B).rank() - 2 * nax;
      pm.assign(perm, perm + rr);
    }
    *out = wrap(ctx, tn::tensordot_t(T(A), T(B), a, b, pm));
  });
}

int tenes_transpose(const tenes_tensor *Tt, const int *axes,
                    tenes_tensor **out) {
  if (!Tt) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(Tt);
  return guard(ctx, [&] {
    TN_REQUIRE(out && axes, "null pointer");
    tn::Axes ax(axes, axes + T(Tt).rank());
    *out = wrap(ctx, tn::transpose(T(Tt), ax));
  });
}

int tenes_svd_top(const tenes_tensor *A, int k, int power_iters,
                  tenes_tensor **U, tenes_tensor **V, double *s, int *k_out) {
  if (!A) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(A);
  return guard(ctx, [&] {
    tn::SvdOptions o;
    o.power_iters = power_iters;
    tn::SvdResult r = tn::svd_top(*ctx->c, T(A), k, o);
    if (U) *U = wrap(ctx, r.U);
    if (V) *V = wrap(ctx, r.V);
    if (s) std::memcpy(s, r.s.data(), r.s.size() * sizeof(double));
    if (k_out) *k_out = (int)r.s.size();
  });
}

int tenes_svd_top_noalloc(tenes_ctx *ctx, const double *A, int m, int n,
                          int k, double *U, double *V, double *s,
                          int *k_out) {
  if (!ctx) return TENES_ERR_ARGUMENT;
  return guard(ctx, [&] {
    tn::noalloc::svd_small(ctx->c->device(), ctx->c->stream(), A, m, n, k, U,
                           V, s, k_out);
  });
}

void tenes_peps_parameters_default(tenes_peps_parameters *p) {
  if (!p) return;
  PEPS_Parameters q;
  p->CHI = q.CHI;
  p->Inverse_projector_cut = q.Inverse_projector_cut;
  p->CTM_Convergence_Epsilon = q.CTM_Convergence_Epsilon;
  p->Max_CTM_Iteration = q.Max_CTM_Iteration;
  p->CTM_Projector_corner = q.CTM_Projector_corner ? 1 : 0;
  p->Use_RSVD = q.Use_RSVD ? 1 : 0;
  p->RSVD_Oversampling_factor = q.RSVD_Oversampling_factor;
  p->print_level = q.print_level;
  p->svd_power_iters = q.svd_power_iters;
  p->rsvd_power_iters = q.rsvd_power_iters;
  p->svd_warm_start = q.svd_warm_start ? 1 : 0;
  p->svd_warm_power_iters = q.svd_warm_power_iters;
  p->contraction_fp32 = q.contraction_fp32;
  p->projector_cut_on_device = q.projector_cut_on_device ? 1 : 0;
  p->graph_sweeps = q.graph_sweeps ? 1 : 0;
}

int tenes_calc_projector_left_block(
    const tenes_tensor *C1, const tenes_tensor *C4, const tenes_tensor *eT1,
    const tenes_tensor *eT6, const tenes_tensor *eT7, const tenes_tensor *eT8,
    const tenes_tensor *Tn1, const tenes_tensor *Tn4,
    const tenes_peps_parameters *p, tenes_tensor **PU, tenes_tensor **PL) {
  if (!C1) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(C1);
  return guard(ctx, [&] {
    TN_REQUIRE(PU && PL, "null output");
    tn::Tensor pu, pl;
    tenes::itps::core::Calc_projector_left_block(
        T(C1), T(C4), T(eT1), T(eT6), T(eT7), T(eT8), T(Tn1), T(Tn4),
        to_params(p), pu, pl);
    *PU = wrap(ctx, pu);
    *PL = wrap(ctx, pl);
  });
}

int tenes_calc_projector_updown_blocks(const tenes_tensor *const *Ts,
                                       const tenes_peps_parameters *p,
                                       tenes_tensor **PU, tenes_tensor **PL) {
  if (!Ts || !Ts[0]) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(Ts[0]);
  return guard(ctx, [&] {
    TN_REQUIRE(PU && PL, "null output");
    tn::Tensor pu, pl;
    tenes::itps::core::Calc_projector_updown_blocks(
        T(Ts[0]), T(Ts[1]), T(Ts[2]), T(Ts[3]), T(Ts[4]), T(Ts[5]), T(Ts[6]),
        T(Ts[7]), T(Ts[8]), T(Ts[9]), T(Ts[10]), T(Ts[11]), T(Ts[12]),
        T(Ts[13]), T(Ts[14]), T(Ts[15]), to_params(p), pu, pl);
    *PU = wrap(ctx, pu);
    *PL = wrap(ctx, pl);
  });
}

int tenes_calc_next_ctm(const tenes_tensor *C1, const tenes_tensor *C4,
                        const tenes_tensor *eT1, const tenes_tensor *eT6,
                        const tenes_tensor *PU, const tenes_tensor *PL,
                        tenes_tensor **C1_out, tenes_tensor **C4_out) {
  if (!C1) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(C1);
  return guard(ctx, [&] {
    TN_REQUIRE(C1_out && C4_out, "null output");
    tn::Tensor a, b;
    tenes::itps::core::Calc_Next_CTM(T(C1), T(C4), T(eT1), T(eT6), T(PU),
                                     T(PL), a, b);
    *C1_out = wrap(ctx, a);
    *C4_out = wrap(ctx, b);
  });
}

int tenes_calc_next_et(const tenes_tensor *eT8, const tenes_tensor *Tn1,
                       const tenes_tensor *PU, const tenes_tensor *PL,
                       tenes_tensor **eT_out) {
  if (!eT8) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(eT8);
  return guard(ctx, [&] {
    TN_REQUIRE(eT_out, "null output");
    tn::Tensor e;
    tenes::itps::core::Calc_Next_eT(T(eT8), T(Tn1), T(PU), T(PL), e);
    *eT_out = wrap(ctx, e);
  });
}

int tenes_env_create(tenes_ctx *ctx, int LX, int LY, int skew,
                     tenes_env **out) {
  return guard(ctx, [&] {
    TN_REQUIRE(out != nullptr, "null output");
    auto *e = new tenes_env;
    e->ctx = ctx;
    e->lat = tenes::itps::SquareLattice(LX, LY, skew);
    for (int k = 0; k < 9; ++k) e->arr[k].resize(e->lat.N_UNIT);
    *out = e;
  });
}

int tenes_env_destroy(tenes_env *env) {
  if (env && env->ctx && env->ctx->c)
    env->ctx->c->forget_svd_starts(
        tenes::itps::environment_id(env->arr[0]));
  delete env;
  return TENES_OK;
}

int tenes_env_set(tenes_env *env, int kind, int site, const tenes_tensor *t) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    TN_REQUIRE(kind >= 0 && kind < 9, "bad tensor kind");
    TN_REQUIRE(site >= 0 && site < env->lat.N_UNIT, "bad site index");
    env->arr[kind][site] = T(t);
  });
}

int tenes_env_get(tenes_env *env, int kind, int site, tenes_tensor **out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    TN_REQUIRE(out != nullptr, "null output");
    TN_REQUIRE(kind >= 0 && kind < 9, "bad tensor kind");
    TN_REQUIRE(site >= 0 && site < env->lat.N_UNIT, "bad site index");
    TN_REQUIRE(env->arr[kind][site].defined(), "tensor not set");
    *out = wrap(env->ctx, env->arr[kind][site]);
  });
}

int tenes_calc_ctm_environment(tenes_env *env, const tenes_peps_parameters *p,
                               int initialize, int *count, double *sig_max) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    for (const auto &t : env->arr[TENES_TN])
      TN_REQUIRE(t.defined(), "site tensor not set");
    double sm = 0.0;
    auto &a = env->arr;
    const int n = tenes::itps::core::Calc_CTM_Environment(
        a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], to_params(p),
        env->lat, initialize != 0, &sm);
    if (count) *count = n;
    if (sig_max) *sig_max = sm;
  });
}

int tenes_calc_ctm_environment_density(tenes_env *env,
                                       const tenes_peps_parameters *p,
                                       int initialize, int *count,
                                       double *sig_max) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    for (const auto &t : env->arr[TENES_TN])
      TN_REQUIRE(t.defined() && t.rank() == 6,
                 "iTPO site tensors (D,D,D,D,d,d) not set");
    auto &a = env->arr;
    // Make_single_tensor_density (ctm_single.cpp:43-50)
    std::vector<tn::Tensor> A(a[8].size());
    for (size_t i = 0; i < A.size(); ++i)
      A[i] = tpo_ket(a[8][i], identity_op(env->ctx, a[8][i].shape[4]));
    const tenes::itps::PEPS_Parameters pp = to_params(p);
    tn::SingleLayer single(*env->ctx->c, true);
    double sm = 0.0;
    const int n = tenes::itps::core::Calc_CTM_Environment(
        a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], A, pp, env->lat,
        initialize != 0, &sm);
    if (count) *count = n;
    if (sig_max) *sig_max = sm;
  });
}

int tenes_batch_calc_ctm_environment(tenes_env *const *envs, int n,
                                     const tenes_peps_parameters *p,
                                     int nparams, int initialize, int *count,
                                     double *sig_max) {
  if (n < 0 || (n > 0 && (!envs || !p)) || !(nparams == 1 || nparams == n))
    return TENES_ERR_ARGUMENT;
  // the environments of one context run in lockstep on one host thread;
  // contexts run concurrently (each has its own streams and allocator)
  std::map<tenes_ctx *, std::vector<int>> groups;
  for (int e = 0; e < n; ++e) {
    if (!envs[e] || !envs[e]->ctx) return TENES_ERR_ARGUMENT;
    groups[envs[e]->ctx].push_back(e);
  }
  std::vector<std::pair<tenes_ctx *, std::vector<int>>> work(groups.begin(),
                                                             groups.end());
  std::vector<int> rc(work.size(), TENES_OK);
  std::vector<std::thread> threads;
  threads.reserve(work.size());
  for (size_t g = 0; g < work.size(); ++g) {
    threads.emplace_back([&, g] {
      tenes_ctx *ctx = work[g].first;
      rc[g] = guard(ctx, [&] {
        TN_HIP_CHECK(hipSetDevice(ctx->c->device()));
        std::vector<tenes::itps::core::CtmEnvironment> list;
        std::vector<tenes::itps::PEPS_Parameters> pps;
        for (int e : work[g].second) {
          tenes_env *env = envs[e];
          for (const auto &t : env->arr[TENES_TN])
            TN_REQUIRE(t.defined(), "site tensor not set");
          auto &a = env->arr;
          list.push_back({{&a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &a[6], &a[7]},
                          &a[8], &env->lat});
          pps.push_back(to_params(&p[nparams == 1 ? 0 : e]));
        }
        std::vector<double> sm;
        const auto c = tenes::itps::core::Calc_CTM_Environment_lockstep(
            list, pps, initialize != 0, &sm);
        for (size_t q = 0; q < c.size(); ++q) {
          const int e = work[g].second[q];
          if (count) count[e] = c[q];
          if (sig_max) sig_max[e] = sm[q];
        }
      });
    });
  }
  for (auto &t : threads) t.join();
  for (size_t g = 0; g < work.size(); ++g)
    if (rc[g] != TENES_OK) {
      g_noctx_err = work[g].first->err;
      return rc[g];
    }
  return TENES_OK;
}

int tenes_env_move(tenes_env *env, int dir, int coord,
                   const tenes_peps_parameters *p) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    auto &a = env->arr;
    const auto q = to_params(p);
    switch (dir) {
      case 0:
        tenes::itps::core::Left_move(a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                                     a[7], a[8], coord, q, env->lat);
        break;
      case 1:
        tenes::itps::core::Right_move(a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                                      a[7], a[8], coord, q, env->lat);
        break;
      case 2:
        tenes::itps::core::Top_move(a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                                    a[7], a[8], coord, q, env->lat);
        break;
      case 3:
        tenes::itps::core::Bottom_move(a[0], a[1], a[2], a[3], a[4], a[5],
                                       a[6], a[7], a[8], coord, q, env->lat);
        break;
      default:
        throw tn::Error("bad move direction", TENES_ERR_ARGUMENT);
    }
  });
}

int tenes_env_save_corners(tenes_env *env) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    for (int k = 0; k < 4; ++k) env->saved[k] = env->arr[k];
  });
}

int tenes_env_check_convergence(tenes_env *env, const tenes_peps_parameters *p,
                                int *converged, double *sig_max) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    for (int k = 0; k < 4; ++k)
      TN_REQUIRE(env->saved[k].size() == env->arr[k].size(),
                 "corners were not saved");
    double sm = 0.0;
    const bool c = tenes::itps::core::Check_Convergence_CTM(
        env->arr[0], env->arr[1], env->arr[2], env->arr[3], env->saved[0],
        env->saved[1], env->saved[2], env->saved[3], to_params(p), env->lat,
        sm);
    if (converged) *converged = c ? 1 : 0;
    if (sig_max) *sig_max = sm;
  });
}

int tenes_env_corner_spectrum(tenes_env *env, int kind, int site, double *out,
                              int *n_out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    TN_REQUIRE(kind >= 0 && kind < 4, "kind must be a corner");
    TN_REQUIRE(site >= 0 && site < env->lat.N_UNIT, "bad site index");
    auto lam = tenes::itps::core::normalized_spectrum(env->arr[kind][site]);
    if (out) std::memcpy(out, lam.data(), lam.size() * sizeof(double));
    if (n_out) *n_out = (int)lam.size();
  });
}

int tenes_env_contract_one_site(tenes_env *env, int site, const double *op,
                                double *out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    TN_REQUIRE(site >= 0 && site < env->lat.N_UNIT, "bad site index");
    TN_REQUIRE(out != nullptr, "null output");
    auto &a = env->arr;
    const int64_t d = a[8][site].shape[4];
    tn::Tensor o = op ? host_op(env->ctx, op, d) : identity_op(env->ctx, d);
    if (is_tpo(env)) {  // Contract_one_site_density_CTM, ctm.cpp:87-112
      const tn::Tensor ket = tpo_ket(a[8][site], o);
      tn::SingleLayer single(*env->ctx->c, true);
      *out = tenes::itps::core::Contract_one_site_iTPS_CTM(
          a[0][site], a[1][site], a[2][site], a[3][site], a[4][site],
          a[5][site], a[6][site], a[7][site], ket, identity_op(env->ctx, 1));
      return;
    }
    *out = tenes::itps::core::Contract_one_site_iTPS_CTM(
        a[0][site], a[1][site], a[2][site], a[3][site], a[4][site], a[5][site],
        a[6][site], a[7][site], a[8][site], o);
  });
}

int tenes_env_contract_two_sites_op12(tenes_env *env, int source, int dx,
                                      int dy, const double *op12,
                                      double *out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    TN_REQUIRE(std::abs(dx) + std::abs(dy) >= 1,
               "op12 contraction needs two distinct sites");
    if (std::abs(dx) > 3 || std::abs(dy) > 3)
      throw tn::Error("op12 contraction: clusters beyond 4 x 4 are not supported",
                      TENES_ERR_UNSUPPORTED);
    TN_REQUIRE(op12 && out, "null pointer");
    TN_REQUIRE(source >= 0 && source < env->lat.N_UNIT, "bad source site");
    auto &a = env->arr;
    Cluster c = cluster_of(env->lat, source, dx, dy);
    if (is_tpo(env) || c.nrow * c.ncol != 2) {
      // the operator split into products of one-site operators by its SVD
      // over (source legs | target legs), twosite_obs.cpp:223-243 (TeNeS's
      // branch for bonds longer than nearest neighbours); each product is
      // one cluster contraction (single-layer for an iTPO)
      const int64_t d1 = a[8][source].shape[4];
      const int64_t d2 = a[8][c.idx[c.target_row][c.target_col]].shape[4];
      std::vector<double> M((size_t)(d1 * d1 * d2 * d2));
      for (int64_t p = 0; p < d1; ++p)
        for (int64_t q = 0; q < d2; ++q)
          for (int64_t r = 0; r < d1; ++r)
            for (int64_t u = 0; u < d2; ++u)
              M[(size_t)((p * d1 + r) * d2 * d2 + q * d2 + u)] =
                  op12[((p * d2 + q) * d1 + r) * d2 + u];
      std::vector<double> U, s, VT;
      tn::host::svd(M, (int)(d1 * d1), (int)(d2 * d2), U, s, VT);
      const int k = (int)s.size();
      double value = 0.0;
      for (int n = 0; n < k; ++n) {
        if (s[n] == 0.0) continue;
        std::vector<double> A((size_t)(d1 * d1)), B((size_t)(d2 * d2));
        for (int64_t i = 0; i < d1 * d1; ++i) A[i] = U[(size_t)(i * k + n)];
        for (int64_t i = 0; i < d2 * d2; ++i)
          B[i] = VT[(size_t)(n * d2 * d2 + i)];
        const tn::Tensor oa = host_op(env->ctx, A.data(), d1);
        const tn::Tensor ob = host_op(env->ctx, B.data(), d2);
        value += s[n] * bond_cluster_value(env, c, &oa, &ob);
      }
      *out = value;
      return;
    }
    if (c.nrow == 2) {
      const int top = c.idx[0][0], bottom = c.idx[1][0];
      const int64_t d1 = a[8][top].shape[4], d2 = a[8][bottom].shape[4];
      tn::Tensor o;
      if (top == source)
        o = tn::from_host(*env->ctx->c, tn::Shape{d1, d2, d1, d2}, op12);
      else
        o = tn::transpose(
            tn::from_host(*env->ctx->c, tn::Shape{d2, d1, d2, d1}, op12),
            tn::Axes{1, 0, 3, 2});
      *out = tenes::itps::core::Contract_two_sites_vertical_op12_iTPS_CTM(
          a[0][top], a[1][top], a[2][bottom], a[3][bottom], a[4][top],
          a[5][top], a[5][bottom], a[6][bottom], a[7][bottom], a[7][top],
          a[8][top], a[8][bottom], o);
    } else {
      const int left = c.idx[0][0], right = c.idx[0][1];
      const int64_t d1 = a[8][left].shape[4], d2 = a[8][right].shape[4];
      tn::Tensor o;
      if (left == source)
        o = tn::from_host(*env->ctx->c, tn::Shape{d1, d2, d1, d2}, op12);
      else
        o = tn::transpose(
            tn::from_host(*env->ctx->c, tn::Shape{d2, d1, d2, d1}, op12),
            tn::Axes{1, 0, 3, 2});
      *out = tenes::itps::core::Contract_two_sites_horizontal_op12_iTPS_CTM(
          a[0][left], a[1][right], a[2][right], a[3][left], a[4][left],
          a[4][right], a[5][right], a[6][right], a[6][left], a[7][left],
          a[8][left], a[8][right], o);
    }
  });
}

int tenes_env_bond_energy(tenes_env *env, int nbonds, const int *source,
                          const int *dx, const int *dy, const double *op12,
                          double *energy, tenes_tensor *const *grad) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    TN_REQUIRE(energy != nullptr, "null output");
    TN_REQUIRE(nbonds >= 0, "negative bond count");
    TN_REQUIRE(nbonds == 0 || (source && dx && dy && op12), "null pointer");
    auto &a = env->arr;
    const int N = env->lat.N_UNIT;
    std::vector<tenes::itps::core::BondTerm> terms(nbonds);
    size_t off = 0;
    for (int b = 0; b < nbonds; ++b) {
      TN_REQUIRE(source[b] >= 0 && source[b] < N, "bad source site");
      TN_REQUIRE(std::abs(dx[b]) + std::abs(dy[b]) == 1,
                 "bond energy needs nearest-neighbour bonds");
      Cluster c = cluster_of(env->lat, source[b], dx[b], dy[b]);
      auto &t = terms[b];
      t.vertical = c.nrow == 2;
      t.s1 = c.idx[0][0];
      t.s2 = t.vertical ? c.idx[1][0] : c.idx[0][1];
      const int64_t d1 = a[8][t.s1].shape[4], d2 = a[8][t.s2].shape[4];
      const double *o = op12 + off;
      off += (size_t)(d1 * d2 * d1 * d2);
      t.op.resize(d1 * d2 * d1 * d2);
      const bool swap = t.s1 != source[b];  // given as (s2, s1, s2', s1')
      for (int64_t p = 0; p < d1; ++p)
        for (int64_t q = 0; q < d2; ++q)
          for (int64_t r = 0; r < d1; ++r)
            for (int64_t u = 0; u < d2; ++u)
              t.op[((p * d2 + q) * d1 + r) * d2 + u] =
                  swap ? o[((q * d1 + p) * d2 + u) * d1 + r]
                       : o[((p * d2 + q) * d1 + r) * d2 + u];
    }
    if (grad) {
      std::vector<tn::Tensor> G(N);
      for (int i = 0; i < N; ++i) {
        TN_REQUIRE(grad[i] != nullptr, "null gradient tensor");
        G[i] = T(grad[i]);
      }
      *energy = tenes::itps::core::bond_energy(a, terms, N, &G);
    } else {
      *energy = tenes::itps::core::bond_energy(a, terms, N, nullptr);
    }
  });
}

int tenes_tensor_to_device(const tenes_tensor *t, double *dst) {
  if (!t || !dst) return TENES_ERR_ARGUMENT;
  return guard(t->owner, [&] {
    const tn::Tensor &x = T(t);
    hipStream_t s = x.ctx()->stream();
    TN_HIP_CHECK(hipMemcpyAsync(dst, x.data(), x.size() * sizeof(double),
                                hipMemcpyDeviceToDevice, s));
    TN_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int tenes_env_contract_multisite(tenes_env *env, int source, int nothers,
                                 const int *dx, const int *dy,
                                 const double *ops, double *out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    TN_REQUIRE(out && ops && (nothers == 0 || (dx && dy)), "null pointer");
    TN_REQUIRE(nothers >= 0 && source >= 0 && source < env->lat.N_UNIT,
               "bad multisite operator");
    auto &a = env->arr;
    // the smallest cluster holding the sites, multisite_obs.cpp:52-155
    int mindx = 0, maxdx = 0, mindy = 0, maxdy = 0;
    for (int i = 0; i < nothers; ++i) {
      mindx = std::min(mindx, dx[i]);
      maxdx = std::max(maxdx, dx[i]);
      mindy = std::min(mindy, dy[i]);
      maxdy = std::max(maxdy, dy[i]);
    }
    const int ncol = maxdx - mindx + 1, nrow = maxdy - mindy + 1;
    const int scol = -mindx, srow = maxdy;
    std::vector<std::vector<int>> idx(nrow, std::vector<int>(ncol));
    std::vector<const tn::Tensor *> C(4), eTt(ncol), eTr(nrow), eTb(ncol), eTl(nrow);
    std::vector<std::vector<const tn::Tensor *>> Tn(nrow, std::vector<const tn::Tensor *>(ncol)),
        op(nrow, std::vector<const tn::Tensor *>(ncol));
    std::vector<tn::Tensor> keep;
    keep.reserve(2 * nrow * ncol + 1);
    for (int row = 0; row < nrow; ++row) {
      for (int col = 0; col < ncol; ++col) {
        idx[row][col] = env->lat.other(source, col - scol, srow - row);
        Tn[row][col] = &a[8][idx[row][col]];
        keep.push_back(identity_op(env->ctx, a[8][idx[row][col]].shape[4]));
        op[row][col] = &keep.back();
      }
      eTl[row] = &a[7][idx[row][0]];
      eTr[row] = &a[5][idx[row][ncol - 1]];
    }
    for (int col = 0; col < ncol; ++col) {
      eTt[col] = &a[4][idx[0][col]];
      eTb[col] = &a[6][idx[nrow - 1][col]];
    }
    C[0] = &a[0][idx[0][0]];
    C[1] = &a[1][idx[0][ncol - 1]];
    C[2] = &a[2][idx[nrow - 1][ncol - 1]];
    C[3] = &a[3][idx[nrow - 1][0]];
    const double norm =
        cluster_value(env, C, eTt, eTr, eTb, eTl, Tn, op);
    size_t off = 0;
    for (int i = 0; i <= nothers; ++i) {
      const int row = i == 0 ? srow : srow - dy[i - 1];
      const int col = i == 0 ? scol : scol + dx[i - 1];
      const int64_t d = a[8][idx[row][col]].shape[4];
      keep.push_back(host_op(env->ctx, ops + off, d));
      off += (size_t)(d * d);
      op[row][col] = &keep.back();
    }
    *out = cluster_value(env, C, eTt, eTr, eTb, eTl, Tn, op) / norm;
  });
}

int tenes_env_correlation(tenes_env *env, int left_site, const double *left_op,
                          int nright, const double *right_ops, int r_max,
                          int direction, double *out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    TN_REQUIRE(left_op && out && (nright == 0 || right_ops), "null pointer");
    TN_REQUIRE(left_site >= 0 && left_site < env->lat.N_UNIT, "bad site index");
    TN_REQUIRE(direction == 0 || direction == 1, "direction must be 0 or 1");
    TN_REQUIRE(r_max >= 0 && nright >= 0, "negative count");
    auto &a = env->arr;
    tn::Context &c = *env->ctx->c;
    const int64_t d = a[8][left_site].shape[4];
    for (int i = 0; i < env->lat.N_UNIT; ++i)
      TN_REQUIRE(a[8][i].shape[4] == d, "one physical dimension for all sites");
    std::vector<double> eye(d * d, 0.0);
    for (int64_t i = 0; i < d; ++i) eye[i * d + i] = 1.0;
    const tn::Tensor I = tn::from_host(c, tn::Shape{d, d}, eye.data());
    const tn::Tensor L = tn::from_host(c, tn::Shape{d, d}, left_op);
    std::vector<tn::Tensor> R(nright);
    for (int j = 0; j < nright; ++j)
      R[j] = tn::from_host(c, tn::Shape{d, d}, right_ops + (size_t)j * d * d);
    namespace K = tenes::itps::core;
    const tn::Axes rot{3, 0, 1, 2, 4};
    // iTPO: the StartCorrelation/Transfer/FinishCorrelation_density_CTM of
    // contract_density_ctm/correlation.cpp:34-95 as the same functions on
    // the single-layer network (the operator traced into the site tensor)
    const bool tpo = is_tpo(env);
    const tn::Tensor one = identity_op(env->ctx, 1);
    auto site = [&](int q, const tn::Tensor &o) {
      const tn::Tensor t = tpo ? tpo_ket(a[8][q], o) : a[8][q];
      return direction == 0 ? t : tn::transpose(t, rot);
    };
    auto opx = [&](const tn::Tensor &o) -> const tn::Tensor & {
      return tpo ? one : o;
    };
    tn::SingleLayer single(c, tpo);
    const int s = left_site;
    tn::Tensor A, An;
    // the column of a vertical correlation is handled as a row of the
    // rotated site tensors (correlation_function.cpp:146-216)
    if (direction == 0) {
      K::StartCorrelation_iTPS_CTM(A, a[0][s], a[3][s], a[4][s], a[6][s], a[7][s], site(s, L), opx(L));
      K::StartCorrelation_iTPS_CTM(An, a[0][s], a[3][s], a[4][s], a[6][s], a[7][s], site(s, I), opx(I));
    } else {
      K::StartCorrelation_iTPS_CTM(A, a[3][s], a[2][s], a[7][s], a[5][s], a[6][s], site(s, L), opx(L));
      K::StartCorrelation_iTPS_CTM(An, a[3][s], a[2][s], a[7][s], a[5][s], a[6][s], site(s, I), opx(I));
    }
    int r_site = s;
    for (int r = 0; r < r_max; ++r) {
      if (direction == 0) {
        r_site = env->lat.right(r_site);
        const int q = r_site;
        const tn::Tensor t = site(q, I);
        const double norm = K::FinishCorrelation_iTPS_CTM(
            An, a[1][q], a[2][q], a[4][q], a[5][q], a[6][q], t, opx(I));
        for (int j = 0; j < nright; ++j)
          out[(size_t)r * nright + j] = K::FinishCorrelation_iTPS_CTM(
              A, a[1][q], a[2][q], a[4][q], a[5][q], a[6][q], site(q, R[j]), opx(R[j])) / norm;
        K::Transfer_iTPS_CTM(A, a[4][q], a[6][q], t);
        K::Transfer_iTPS_CTM(An, a[4][q], a[6][q], t);
      } else {
        r_site = env->lat.top(r_site);
        const int q = r_site;
        const tn::Tensor t = site(q, I);
        const double norm = K::FinishCorrelation_iTPS_CTM(
            An, a[0][q], a[1][q], a[7][q], a[4][q], a[5][q], t, opx(I));
        for (int j = 0; j < nright; ++j)
          out[(size_t)r * nright + j] = K::FinishCorrelation_iTPS_CTM(
              A, a[0][q], a[1][q], a[7][q], a[4][q], a[5][q], site(q, R[j]), opx(R[j])) / norm;
        K::Transfer_iTPS_CTM(A, a[7][q], a[5][q], t);
        K::Transfer_iTPS_CTM(An, a[7][q], a[5][q], t);
      }
    }
  });
}

int tenes_env_contract_cluster(tenes_env *env, int source, int dx, int dy,
                               const double *op_source,
                               const double *op_target, double *out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    TN_REQUIRE(out != nullptr, "null output");
    TN_REQUIRE(source >= 0 && source < env->lat.N_UNIT, "bad source site");
    auto &a = env->arr;
    Cluster c = cluster_of(env->lat, source, dx, dy);
    tn::Tensor os, ot;
    if (op_source)
      os = host_op(env->ctx, op_source,
                   a[8][c.idx[c.source_row][c.source_col]].shape[4]);
    if (op_target)
      ot = host_op(env->ctx, op_target,
                   a[8][c.idx[c.target_row][c.target_col]].shape[4]);
    *out = bond_cluster_value(env, c, op_source ? &os : nullptr,
                              op_target ? &ot : nullptr);
  });
}

int tenes_create_environment_two_sites(const tenes_tensor *const *Ts,
                                       tenes_tensor **out) {
  if (!Ts || !Ts[0]) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(Ts[0]);
  return guard(ctx, [&] {
    TN_REQUIRE(out != nullptr, "null output");
    *out = wrap(ctx, tenes::itps::core::Create_Environment_two_sites(
                         T(Ts[0]), T(Ts[1]), T(Ts[2]), T(Ts[3]), T(Ts[4]),
                         T(Ts[5]), T(Ts[6]), T(Ts[7]), T(Ts[8]), T(Ts[9]),
                         T(Ts[10]), T(Ts[11])));
  });
}

namespace {

tenes::itps::FU_Parameters to_fu(const tenes_fu_parameters *p) {
  tenes::itps::FU_Parameters q;
  if (p) {
    q.Inverse_Env_cut = p->Inverse_Env_cut;
    q.Full_Inverse_precision = p->Full_Inverse_precision;
    q.Full_Convergence_Epsilon = p->Full_Convergence_Epsilon;
    q.Full_max_iteration = p->Full_max_iteration;
    q.Full_Gauge_Fix = p->Full_Gauge_Fix != 0;
    q.Full_Use_FastFullUpdate = p->Full_Use_FastFullUpdate != 0;
  }
  return q;
}

}  // namespace

void tenes_fu_parameters_default(tenes_fu_parameters *p) {
  if (!p) return;
  const tenes::itps::FU_Parameters q;
  p->Inverse_Env_cut = q.Inverse_Env_cut;
  p->Full_Inverse_precision = q.Full_Inverse_precision;
  p->Full_Convergence_Epsilon = q.Full_Convergence_Epsilon;
  p->Full_max_iteration = q.Full_max_iteration;
  p->Full_Gauge_Fix = q.Full_Gauge_Fix ? 1 : 0;
  p->Full_Use_FastFullUpdate = q.Full_Use_FastFullUpdate ? 1 : 0;
}

int tenes_full_update_bond(const tenes_tensor *const *Ts, const double *op12,
                           int connect1, const tenes_fu_parameters *fp,
                           tenes_tensor **Tn1_new, tenes_tensor **Tn2_new) {
  if (!Ts || !Ts[0]) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(Ts[0]);
  return guard(ctx, [&] {
    TN_REQUIRE(op12 && Tn1_new && Tn2_new, "null pointer");
    for (int i = 0; i < 12; ++i) TN_REQUIRE(Ts[i] != nullptr, "null tensor handle");
    const tn::Tensor &A = T(Ts[10]), &B = T(Ts[11]);
    TN_REQUIRE(A.rank() == 5 && B.rank() == 5, "site tensors must have 5 legs");
    tn::host::HT op({A.shape[4], B.shape[4], A.shape[4], B.shape[4]});
    std::copy(op12, op12 + op.size(), op.v.begin());
    tn::Tensor n1, n2;
    tenes::itps::core::Full_update_bond(
        T(Ts[0]), T(Ts[1]), T(Ts[2]), T(Ts[3]), T(Ts[4]), T(Ts[5]), T(Ts[6]),
        T(Ts[7]), T(Ts[8]), T(Ts[9]), A, B, op, connect1, to_fu(fp), n1, n2);
    *Tn1_new = wrap(ctx, n1);
    *Tn2_new = wrap(ctx, n2);
  });
}

int tenes_env_full_update(tenes_env *env, int nops, const int *source,
                          const int *source_leg, const int *op_dims,
                          const double *ops, int nsteps,
                          const tenes_peps_parameters *pp,
                          const tenes_fu_parameters *fp) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    for (const auto &t : env->arr[TENES_TN])
      TN_REQUIRE(t.defined(), "site tensor not set");
    TN_REQUIRE(nops >= 0 && (nops == 0 || (source && source_leg && ops)),
               "null pointer");
    auto &a = env->arr;
    const int N = env->lat.N_UNIT;
    std::vector<tn::host::HT> op(nops);
    size_t off = 0;
    for (int k = 0; k < nops; ++k) {
      TN_REQUIRE(source[k] >= 0 && source[k] < N, "bad source site");
      TN_REQUIRE(source_leg[k] >= 0 && source_leg[k] < 4, "bad source leg");
      std::vector<int64_t> dims(4);
      if (op_dims) {
        for (int q = 0; q < 4; ++q) {
          TN_REQUIRE(op_dims[4 * k + q] >= 1, "bad operator dimension");
          dims[q] = op_dims[4 * k + q];
        }
      } else {
        const int t = env->lat.neighbor(source[k], source_leg[k]);
        const int64_t ds = a[8][source[k]].shape[4], dt = a[8][t].shape[4];
        dims = {ds, dt, ds, dt};
      }
      op[k] = tn::host::HT(dims);
      std::copy(ops + off, ops + off + op[k].size(), op[k].v.begin());
      off += (size_t)op[k].size();
    }
    if (nsteps <= 0) return;
    const auto q = to_params(pp);
    const auto f = to_fu(fp);
    tenes::itps::core::Calc_CTM_Environment(a[0], a[1], a[2], a[3], a[4], a[5],
                                            a[6], a[7], a[8], q, env->lat,
                                            true, nullptr);
    for (int step = 0; step < nsteps; ++step)
      for (int k = 0; k < nops; ++k)
        tenes::itps::full_update_step(a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                                      a[7], a[8], source[k], source_leg[k],
                                      op[k], q, f, env->lat);
  });
}

int tenes_simple_update_bond(const tenes_tensor *Tn1, const tenes_tensor *Tn2,
                             const double *lambda1, const double *lambda2,
                             const tenes_tensor *op12, int connect1,
                             double inverse_lambda_cut, tenes_tensor **Tn1_new,
                             tenes_tensor **Tn2_new, double *lambda_c) {
  if (!Tn1) return TENES_ERR_ARGUMENT;
  tenes_ctx *ctx = ctx_of(Tn1);
  return guard(ctx, [&] {
    TN_REQUIRE(lambda1 && lambda2 && Tn1_new && Tn2_new && lambda_c,
               "null pointer");
    TN_REQUIRE(T(Tn1).rank() == 5 && T(Tn2).rank() == 5,
               "site tensors must have 5 legs");
    std::vector<std::vector<double>> l1(4), l2(4);
    int64_t o1 = 0, o2 = 0;
    for (int l = 0; l < 4; ++l) {
      l1[l].assign(lambda1 + o1, lambda1 + o1 + T(Tn1).shape[l]);
      o1 += T(Tn1).shape[l];
      l2[l].assign(lambda2 + o2, lambda2 + o2 + T(Tn2).shape[l]);
      o2 += T(Tn2).shape[l];
    }
    tn::Tensor a, b;
    std::vector<double> lc;
    tenes::itps::core::Simple_update_bond(T(Tn1), T(Tn2), l1, l2, T(op12),
                                          connect1, inverse_lambda_cut, a, b,
                                          lc);
    *Tn1_new = wrap(ctx, a);
    *Tn2_new = wrap(ctx, b);
    std::memcpy(lambda_c, lc.data(), lc.size() * sizeof(double));
  });
}

static void ensure_lambda(tenes_env *env) {
  const int N = env->lat.N_UNIT;
  if ((int)env->lambda.size() == N) return;
  env->lambda.assign(N, std::vector<std::vector<double>>(4));
  for (int i = 0; i < N; ++i) {
    TN_REQUIRE(env->arr[TENES_TN][i].defined(), "site tensor not set");
    for (int l = 0; l < 4; ++l)
      env->lambda[i][l].assign(env->arr[TENES_TN][i].shape[l], 1.0);
  }
}

int tenes_env_simple_update(tenes_env *env, int nops, const int *source_sites,
                            const int *source_legs,
                            const tenes_tensor *const *ops, int nsteps,
                            double inverse_lambda_cut) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    TN_REQUIRE(nops >= 0 && (nops == 0 || (source_sites && source_legs && ops)),
               "null pointer");
    ensure_lambda(env);
    std::vector<tenes::itps::core::EvolutionOp> up;
    for (int k = 0; k < nops; ++k) {
      TN_REQUIRE(source_sites[k] >= 0 && source_sites[k] < env->lat.N_UNIT,
                 "bad source site");
      TN_REQUIRE(source_legs[k] >= 0 && source_legs[k] < 4, "bad source leg");
      up.push_back({source_sites[k], source_legs[k], T(ops[k])});
    }
    if (!is_tpo(env)) {
      tenes::itps::core::Simple_update_steps(env->arr[TENES_TN], env->lambda,
                                             env->lat, up, nsteps,
                                             inverse_lambda_cut);
      return;
    }
    // iTPO (simple_update.cpp:141-178): the two physical legs merged
    // (ket index the slower) and the operator applied as kron(U, conj(U))
    auto &Tn = env->arr[TENES_TN];
    std::vector<tn::Tensor> M(Tn.size());
    for (size_t i = 0; i < Tn.size(); ++i) {
      const auto &t = Tn[i];
      M[i] = tn::reshape(t, tn::Shape{t.shape[0], t.shape[1], t.shape[2],
                                      t.shape[3], t.shape[4] * t.shape[5]});
    }
    for (auto &u : up) {
      TN_REQUIRE(u.op.rank() == 4, "iTPO evolution operators must be two-site");
      std::vector<double> h(u.op.size());
      tn::to_host(u.op, h.data());
      const auto &sh = u.op.shape;
      tn::Shape ks{sh[0] * sh[0], sh[1] * sh[1], sh[2] * sh[2], sh[3] * sh[3]};
      std::vector<double> k((size_t)(ks[0] * ks[1] * ks[2] * ks[3]));
      for (int64_t i0 = 0; i0 < sh[0] * sh[0]; ++i0)
        for (int64_t i1 = 0; i1 < sh[1] * sh[1]; ++i1)
          for (int64_t i2 = 0; i2 < sh[2] * sh[2]; ++i2)
            for (int64_t i3 = 0; i3 < sh[3] * sh[3]; ++i3) {
              const int64_t a0 = i0 / sh[0], b0 = i0 % sh[0];
              const int64_t a1 = i1 / sh[1], b1 = i1 % sh[1];
              const int64_t a2 = i2 / sh[2], b2 = i2 % sh[2];
              const int64_t a3 = i3 / sh[3], b3 = i3 % sh[3];
              const double ua = h[((a0 * sh[1] + a1) * sh[2] + a2) * sh[3] + a3];
              const double ub = h[((b0 * sh[1] + b1) * sh[2] + b2) * sh[3] + b3];
              k[((i0 * ks[1] + i1) * ks[2] + i2) * ks[3] + i3] = ua * ub;
            }
      u.op = tn::from_host(*env->ctx->c, ks, k.data());
    }
    tenes::itps::core::Simple_update_steps(M, env->lambda, env->lat, up,
                                           nsteps, inverse_lambda_cut);
    for (size_t i = 0; i < Tn.size(); ++i) {
      const auto &t = M[i];
      const int64_t d = Tn[i].shape[4];
      Tn[i] = tn::reshape(t, tn::Shape{t.shape[0], t.shape[1], t.shape[2],
                                       t.shape[3], d, d});
    }
  });
}

int tenes_env_get_lambda(tenes_env *env, int site, int leg, double *out,
                         int *n_out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    TN_REQUIRE(site >= 0 && site < env->lat.N_UNIT && leg >= 0 && leg < 4,
               "bad site or leg");
    ensure_lambda(env);
    const auto &v = env->lambda[site][leg];
    if (out) std::memcpy(out, v.data(), v.size() * sizeof(double));
    if (n_out) *n_out = (int)v.size();
  });
}

int tenes_env_save_tensors(tenes_env *env, const char *dir, int chi) {
  if (!env || !dir) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    ensure_lambda(env);
    tenes::itps::CheckpointState st{env->arr, &env->lambda, env->lat.N_UNIT};
    tenes::itps::save_tensors(st, dir, chi);
  });
}

int tenes_env_load_tensors(tenes_env *env, const char *dir, int chi) {
  if (!env || !dir) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    tenes::itps::CheckpointState st{env->arr, &env->lambda, env->lat.N_UNIT};
    tenes::itps::load_tensors(st, *env->ctx->c, dir, chi);
  });
}

void tenes_tm_parameters_default(tenes_tm_parameters *p) {
  if (!p) return;
  tenes::itps::TransferMatrix_Parameters q;
  p->num_eigvals = q.num_eigvals;
  p->maxdim_dense_eigensolver = q.maxdim_dense_eigensolver;
  p->arnoldi_maxdim = q.arnoldi_maxdim;
  p->arnoldi_restartdim = q.arnoldi_restartdim;
  p->arnoldi_maxiter = q.arnoldi_maxiter;
  p->arnoldi_rtol = q.arnoldi_rtol;
}

int tenes_env_transfer_matrix_eigenvalues(tenes_env *env, int dir,
                                          int fixed_coord,
                                          const tenes_tm_parameters *p,
                                          double *eig, int *n_out) {
  if (!env) return TENES_ERR_ARGUMENT;
  return guard(env->ctx, [&] {
    check_env(env);
    TN_REQUIRE(dir == 0 || dir == 1, "dir must be 0 (x) or 1 (y)");
    TN_REQUIRE(eig != nullptr, "null output");
    tenes::itps::TransferMatrix_Parameters q;
    if (p) {
      q.num_eigvals = p->num_eigvals;
      q.maxdim_dense_eigensolver = p->maxdim_dense_eigensolver;
      q.arnoldi_maxdim = p->arnoldi_maxdim;
      q.arnoldi_restartdim = p->arnoldi_restartdim;
      q.arnoldi_maxiter = p->arnoldi_maxiter;
      q.arnoldi_rtol = p->arnoldi_rtol;
    }
    TN_REQUIRE(q.num_eigvals >= 1, "num_eigvals must be positive");
    auto &a = env->arr;
    tenes::itps::TransferMatrix_ctm tm(env->lat, a[8], a[0], a[1], a[2], a[3],
                                       a[7], a[4], a[5], a[6]);
    auto ev = tm.eigenvalues(dir, fixed_coord, q);
    for (size_t i = 0; i < ev.size(); ++i) {
      eig[2 * i] = ev[i].real();
      eig[2 * i + 1] = ev[i].imag();
    }
    if (n_out) *n_out = (int)ev.size();
  });
}

}  // extern "C"
