// CTMRG on the GPU: projectors, renormalisation, directional moves,
// convergence check and the environment driver.  Each function follows the
// contraction order of the TeNeS function it replaces (cited per function);
// every tensordot runs as one gather-GEMM launch on the context stream.
#include "ctm.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>

#include "kernels.hpp"

namespace tenes {
namespace itps {

using tn::Axes;
using tn::Shape;

// ---------------------------------------------------------------------------
// SquareLattice (src/SquareLattice.cpp:48-137)
// ---------------------------------------------------------------------------
static int gcd_i(int a, int b) {
  while (b) {
    int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

SquareLattice::SquareLattice(int X, int Y, int skew_in) {
  TN_REQUIRE(X > 0 && Y > 0, "SquareLattice: LX and LY must be positive");
  LX = X;
  LY = Y;
  N_UNIT = LX * LY;
  skew = ((skew_in % LX) + LX) % LX;
  LX_noskew = LX;
  LY_noskew = skew == 0 ? LY : LY * ((LX / gcd_i(LX, skew) * skew) / skew);
  N_UNIT_noskew = LX_noskew * LY_noskew;
  nn_.resize(N_UNIT);
  for (int iy = 0; iy < LY; ++iy)
    for (int ix = 0; ix < LX; ++ix) {
      const int i = ix + iy * LX;
      nn_[i][0] = ((ix - 1 + LX) % LX) + iy * LX;
      nn_[i][1] = ix + ((iy + 1) % LY) * LX;
      nn_[i][2] = ((ix + 1) % LX) + iy * LX;
      nn_[i][3] = ix + ((iy - 1 + LY) % LY) * LX;
    }
  if (skew != 0) {
    for (int ix = 0; ix < LX; ++ix) {
      nn_[ix][3] = ((ix + LX + skew) % LX) + (LY - 1) * LX;
      nn_[N_UNIT + ix - LX][1] = ((ix + LX - skew) % LX);
    }
  }
  parity_.assign(N_UNIT, 0);
  for (int ix = 0; ix < LX; ++ix) parity_[ix] = (ix % 2 != 0) ? -1 : 1;
  for (int iy = 1; iy < LY; ++iy)
    for (int ix = 0; ix < LX; ++ix) {
      const int site = index(ix, iy);
      parity_[site] = -parity_[bottom(site)];
    }
}

int SquareLattice::index(int x, int y) const {
  int y_offset = 0;
  if (y >= LY)
    y_offset = y / LY;
  else if (y < 0)
    y_offset = (y + 1) / LY - 1;
  y -= LY * y_offset;
  x -= skew * y_offset;
  x %= LX;
  if (x < 0) x += LX;
  return x + y * LX;
}

int SquareLattice::other(int idx, int dx, int dy) const {
  while (dx != 0) {
    if (dx > 0) {
      idx = right(idx);
      --dx;
    } else {
      idx = left(idx);
      ++dx;
    }
  }
  while (dy != 0) {
    if (dy > 0) {
      idx = top(idx);
      --dy;
    } else {
      idx = bottom(idx);
      ++dy;
    }
  }
  return idx;
}

namespace core {

namespace {

tn::SvdOptions svd_opts(const PEPS_Parameters &p, bool rsvd) {
  tn::SvdOptions o;
  o.power_iters = rsvd ? p.rsvd_power_iters : p.svd_power_iters;
  return o;
}

// Truncation of ctm.cpp:275-288: keep s[i]/s[0] > cut, at most `cut` values;
// returns the kept count and fills s_c = 1/sqrt(s).
int truncate_spectrum(const std::vector<double> &s, int cut, double eps,
                      std::vector<double> &s_c) {
  s_c.clear();
  if (s.empty()) return 0;
  const double denom = s[0];
  for (int i = 0; i < cut; ++i) {
    if (s[i] / denom > eps) {
      s_c.push_back(1.0 / std::sqrt(s[i]));
    } else {
      cut = i;
      break;
    }
  }
  return cut;
}

tensor identity_projector(tn::Context &ctx, int64_t e78, int64_t t41) {
  tensor I = tn::empty(ctx, Shape{e78, e78});
  tn::kern::eye(I.data(), e78, ctx.stream());
  return tn::reshape(I, Shape{e78, t41, t41, e78});
}

// The two contractions of a block tensor with the ket and the bra site
// tensor that follow each other in the enlarged corners (ctm.cpp:209-212,
// 224-227) and in Calc_Next_eT (ctm.cpp:532-536), as one kernel
// (kern::site_apply).  legA, legB are the legs of Tn (l, u, r, d, p) met by
// the block tensor's ket legs kA, kB (its bra legs bA, bB meet the same
// legs of the bra); the result keeps Tn's other two of (l, u, r, d) in
// increasing order.  The fused kernel applies to FP64 double-layer networks
// with all four bond dimensions equal to D <= 5 and d = 2; elsewhere the
// callers run the two contractions.
bool fused_site_ok(const tensor &Tn) {
  tn::Context &ctx = *Tn.ctx();
  if (!ctx.fused_site() || ctx.contraction_mode() != 0 || ctx.single_layer() || Tn.rank() != 5)
    return false;
  const int64_t D = Tn.shape[0];
  for (int l = 1; l < 4; ++l)
    if (Tn.shape[l] != D) return false;
  return tn::kern::site_apply_supported((int)D, (int)Tn.shape[4]);
}

// X block-major (s, f, kA, kB, bA, bB); result corner layout (oA, s, boA,
// oB, f, boB) or block-major (s, f, oA, oB, boA, boB) (edge).
tensor fused_site(const tensor &X, const tensor &Tn, int legA, int legB, bool edge) {
  tn::Context &ctx = *X.ctx();
  int o[2], n = 0;
  for (int l = 0; l < 4; ++l)
    if (l != legA && l != legB) o[n++] = l;
  const int D = (int)Tn.shape[0];
  TN_REQUIRE(X.rank() == 6 && X.shape[2] == D && X.shape[3] == D && X.shape[4] == D &&
                 X.shape[5] == D,
             "fused site contraction: block tensor shape");
  tn::kern::SiteArgs a;
  a.Ds = (int)X.shape[0];
  a.Df = (int)X.shape[1];
  a.D = D;
  a.p = (int)Tn.shape[4];
  // operand arrays (real tensors: the bra is the ket)
  const tensor TK = tn::transpose(Tn, Axes{legA, legB, 4, o[0], o[1]});
  const tensor TB = tn::transpose(Tn, Axes{legA, legB, o[0], o[1], 4});
  tensor out = edge ? tn::empty(ctx, Shape{a.Ds, a.Df, D, D, D, D})
                    : tn::empty(ctx, Shape{D, a.Ds, D, D, a.Df, D});
  a.X = X.data();
  a.TK = TK.data();
  a.TB = TB.data();
  a.out = out.data();
  tn::kern::site_apply(a, edge, ctx.stream());
  return out;
}

// Enlarged corner of ctm.cpp:207-212 (and its rotated uses):
// LT = Tn1*(((C1*eT1)*eT8)*Tn1c), reference legs (t12, t41, e12, e78, tc12,
// tc41).  Stored with the legs grouped as a matrix,
// (t12, e12, tc12 | t41, e78, tc41) = reference transpose (0,2,4,1,3,5),
// fused into the last contraction's store: the projector GEMMs that consume
// it then read plain row-major operands.
tensor enlarged_LT(const tensor &C1, const tensor &eT1, const tensor &eT8,
                   const tensor &Tn1) {
  const tensor T1 = tn::tensordot(C1, eT1, Axes{1}, Axes{0});
  if (fused_site_ok(Tn1))  // (c, e, k1, k8, b1, b8): eT1's legs meet Tn1's u, eT8's its l
    return fused_site(tn::tensordot_t(T1, eT8, Axes{0}, Axes{1}, Axes{0, 3, 1, 4, 2, 5}), Tn1, 1, 0,
                      false);
  const tensor T2 = tn::tensordot(T1, eT8, Axes{0}, Axes{1});  // (c, k1, b1, e, k8, b8)
  return tn::tensordot_t(Tn1, tn::tensordot(T2, tn::bra(Tn1), Axes{2, 5}, Axes{1, 0}),
                         Axes{0, 1, 4}, Axes{3, 1, 6}, Axes{0, 2, 4, 1, 3, 5});
}
// ctm.cpp:222-227: LB = Tn4*(((C4*eT7)*eT6)*Tn4c), reference legs (t41,
// t34, e78, e56, tc41, tc34); stored as (t41, e78, tc41 | t34, e56, tc34).
tensor enlarged_LB(const tensor &C4, const tensor &eT7, const tensor &eT6,
                   const tensor &Tn4) {
  const tensor T1 = tn::tensordot(C4, eT7, Axes{1}, Axes{0});
  if (fused_site_ok(Tn4))  // (c, e, k7, k6, b7, b6): eT7's legs meet Tn4's l, eT6's its d
    return fused_site(tn::tensordot_t(T1, eT6, Axes{0}, Axes{1}, Axes{0, 3, 1, 4, 2, 5}), Tn4, 0, 3,
                      false);
  const tensor T2 = tn::tensordot(T1, eT6, Axes{0}, Axes{1});  // (c, k7, b7, e, k6, b6)
  return tn::tensordot_t(Tn4, tn::tensordot(T2, tn::bra(Tn4), Axes{2, 5}, Axes{0, 3}),
                         Axes{0, 3, 4}, Axes{1, 3, 6}, Axes{0, 2, 4, 1, 3, 5});
}

// One projector computation, split at its only host synchronisation (the
// singular values decide the cut): *_begin issues the enlarged corners, R
// and the truncated SVD on the current stream, projectors_finish waits for
// the singular values and issues the projector GEMMs.
struct ProjectorJob {
  bool done = false;  // identity projector (t41 == 1): PU, PL already set
  // cut decided on the device (fixed shapes, no host wait), else on the host
  bool device_cut = false;
  tensor PU, PL;
  tensor Ablk, Bblk;
  Axes axA, axB;
  // Matrix-free projector matrix (the Use_RSVD branch): PL = finish_a(U)
  // and PU = finish_b(V) apply the blocks of R to the (r0, r1, rb, k) /
  // (c0, c1, cb, k) singular vectors; else Ablk, Bblk with axA, axB.
  std::function<tensor(const tensor &)> finish_a, finish_b;
  int64_t r0 = 0, r1 = 0, rb = 0, c0 = 0, c1 = 0, cb = 0;  // (t, e, tc) legs
  int kmax = 0;
  // > 0: the truncation of the single-layer up-down projectors' full-SVD
  // branch (ctm_single.cpp:396-409), whose loop runs over the first e78
  // singular values (this value) instead of the first cut
  int e78_loop = 0;
  int slot = 0;  // pinned slot of the singular values
  tn::SvdPending sv;
  // warm-start entry to update with the Ritz-value range once known
  bool warm_store = false;
  std::array<int64_t, 5> wkey{};
};

constexpr int kSyncSlot = tn::Context::kSlotProjectorSync;

// Context contraction modes (0 FP64, 1 FP32, 2 split bf16) of the moves'
// contractions and of the projector matrix R = LT LB for
// PEPS_Parameters::contraction_fp32 = 0, 1, 2, 3.
int move_contraction_mode(const PEPS_Parameters &pp) {
  return pp.contraction_fp32 == 3 ? 2 : (pp.contraction_fp32 != 0 ? 1 : 0);
}
int r_contraction_mode(const PEPS_Parameters &pp) {
  return pp.contraction_fp32 == 3 ? 2 : (pp.contraction_fp32 == 2 ? 1 : 0);
}

// From the SVD of the (rows x cols) matrix `Rm`, build the oblique projector
// pair: PU = Bblk * (VT_c^dagger s^-1/2), PL = Ablk * (U_c^* s^-1/2), both
// transposed to (e_in, t, tc, cut) (ctm.cpp:290-301 / 466-475).
// Projector identity (environment, move, coordinate, row) for the
// warm-started range finder; nullptr: no warm start.
using ProjKey = std::array<int64_t, 4>;

void projectors_begin(tn::Context &ctx, const tensor &Rm, int64_t r0,
                      int64_t r1, int64_t rb, int64_t c0, int64_t c1,
                      int64_t cb, int chi_cap, const tensor &Ablk,
                      const Axes &axA, const tensor &Bblk, const Axes &axB,
                      const PEPS_Parameters &pp, bool rsvd, int slot,
                      const ProjKey *key, ProjectorJob &job,
                      const tn::LinearOp *Rop = nullptr) {
  const int64_t nr = r0 * r1 * rb, nc = c0 * c1 * cb;
  job.kmax = (int)std::min<int64_t>(std::min<int64_t>(nr, nc), chi_cap);
  tn::SvdOptions opts = svd_opts(pp, rsvd);
  const bool warm = key != nullptr && pp.svd_warm_start && !rsvd;
  std::array<int64_t, 5> wkey{};
  if (warm) {
    wkey = {(*key)[0], (*key)[1], (*key)[2], (*key)[3], nc};
    auto it = ctx.svd_start.find(wkey);
    if (it != ctx.svd_start.end()) {
      opts.start = &it->second.V;
      opts.start_sigma = &it->second.s;
      opts.start_range = it->second.range;
      opts.start_power_iters = pp.svd_warm_power_iters;
    }
  }
  job.sv = Rop ? tn::svd_top_begin(ctx, *Rop, job.kmax, opts, slot)
              : tn::svd_top_begin(ctx, tn::reshape(Rm, Shape{nr, nc}), job.kmax, opts, slot);
  job.warm_store = warm && job.sv.r.Vbasis.defined();
  if (job.warm_store) {
    ctx.svd_start[wkey] = {job.sv.r.Vbasis, job.sv.r.sbasis, 0.0};
    job.wkey = wkey;
  }
  job.slot = slot;
  job.Ablk = Ablk;
  job.Bblk = Bblk;
  job.axA = axA;
  job.axB = axB;
  job.r0 = r0;
  job.r1 = r1;
  job.rb = rb;
  job.c0 = c0;
  job.c1 = c1;
  job.cb = cb;
}

// PU = Bblk V, PL = Ablk U (ctm.cpp:290-301), legs (e_in, t, tc, cut);
// the blocks are released afterwards.
void projector_products(ProjectorJob &job, const tensor &Uc, const tensor &Vc) {
  if (job.finish_a) {
    job.PU = job.finish_b(Vc);
    job.PL = job.finish_a(Uc);
  } else {
    job.PU = tn::tensordot_t(job.Bblk, Vc, job.axB, Axes{0, 1, 2}, Axes{1, 0, 2, 3});
    job.PL = tn::tensordot_t(job.Ablk, Uc, job.axA, Axes{0, 1, 2}, Axes{1, 0, 2, 3});
  }
  job.Ablk = tensor();
  job.Bblk = tensor();
  job.finish_a = nullptr;
  job.finish_b = nullptr;
}

void projectors_finish(tn::Context &ctx, ProjectorJob &job,
                       const PEPS_Parameters &pp) {
  if (job.done) return;
  if (job.device_cut) {
    // fixed shape: all k = min(rows, cols, CHI) columns, the dropped ones
    // scaled by zero (ctm.cpp:281-288 decided by projector_scale), applied
    // to the projector GEMMs' outputs: B (V s^-1/2) = (B V) s^-1/2
    const int nall = job.sv.nall;
    tn::SvdResult sv = tn::svd_top_detach(job.sv);
    ctx.pending_svd.push_back({job.wkey, job.warm_store, job.slot, nall});
    const int64_t k = sv.s_dev.size();
    tensor sc = job.e78_loop > 0 ? tn::zeros(ctx, Shape{k}) : tn::empty(ctx, Shape{k});
    // the single-layer up-down rule scales only the first e78 directions
    // (the rest stay zero columns, as the reference's zero-filled s_c)
    const int64_t ks = job.e78_loop > 0 ? std::min<int64_t>(k, job.e78_loop) : k;
    tn::kern::projector_scale(sv.s_dev.data(), (int)ks, pp.Inverse_projector_cut,
                              sc.data(), ctx.stream());
    tensor Uc = tn::reshape(sv.U, Shape{job.r0, job.r1, job.rb, k});
    tensor Vc = tn::reshape(sv.V, Shape{job.c0, job.c1, job.cb, k});
    projector_products(job, Uc, Vc);
    tn::multiply_vector_dev(job.PU, sc.data(), 3);
    tn::multiply_vector_dev(job.PL, sc.data(), 3);
    job.done = true;
    return;
  }
  tn::SvdResult sv = tn::svd_top_end(job.sv);
  if (job.warm_store && !sv.s_all.empty()) {
    auto it = ctx.svd_start.find(job.wkey);
    if (it != ctx.svd_start.end())
      it->second.range = sv.s_all.back() > 0.0
                             ? sv.s_all.front() / sv.s_all.back()
                             : 0.0;
  }
  std::vector<double> s_c;
  const int cut0 = std::min<int>(job.kmax, (int)sv.s.size());
  int cut = 0;
  if (job.e78_loop > 0) {
    // ctm_single.cpp:396-415: s_c starts as cut0 zeros; the loop over the
    // first e78 values stops at the first one at or below the cutoff and
    // makes that index the cut, which can exceed cut0: entries from cut0 on
    // are zero (the vector's resize), and so are entries from e78 on
    const std::vector<double> &sa = sv.s_all.size() > sv.s.size() ? sv.s_all : sv.s;
    const int lim = std::min<int>(job.e78_loop, (int)sa.size());
    cut = cut0;
    for (int i = 0; i < lim; ++i)
      if (!(sa[i] / sa[0] > pp.Inverse_projector_cut)) {
        cut = i;
        break;
      }
    s_c.assign(std::max(cut, 0), 0.0);
    for (int j = 0; j < std::min(std::min(cut, cut0), job.e78_loop); ++j)
      s_c[j] = 1.0 / std::sqrt(sa[j]);
  } else {
    cut = truncate_spectrum(sv.s, cut0, pp.Inverse_projector_cut, s_c);
  }
  TN_REQUIRE(cut >= 1, "projector: no singular value above the cutoff");
  tensor sc = tn::from_host(ctx, Shape{cut}, s_c.data());
  // a cut beyond the computed triplets (single-layer up-down rule) adds
  // zero-scaled columns: zero columns
  const int kc = std::min<int>(cut, (int)sv.U.shape[1]);
  tensor Uc = tn::slice(sv.U, 1, 0, kc);
  tensor Vc = tn::slice(sv.V, 1, 0, kc);
  if (kc < cut) {
    Uc = tn::extend(Uc, Shape{Uc.shape[0], cut});
    Vc = tn::extend(Vc, Shape{Vc.shape[0], cut});
  }
  tn::multiply_vector_dev(Uc, sc.data(), 1);
  tn::multiply_vector_dev(Vc, sc.data(), 1);
  Uc = tn::reshape(Uc, Shape{job.r0, job.r1, job.rb, cut});
  Vc = tn::reshape(Vc, Shape{job.c0, job.c1, job.cb, cut});
  projector_products(job, Uc, Vc);
  job.done = true;
}

// ctm.cpp:174-318 up to the SVD
void left_block_begin(const tensor &C1, const tensor &C4, const tensor &eT1,
                      const tensor &eT6, const tensor &eT7, const tensor &eT8,
                      const tensor &Tn1, const tensor &Tn4,
                      const PEPS_Parameters &pp, int slot, const ProjKey *key,
                      ProjectorJob &job) {
  tn::Context &ctx = *C1.ctx();
  const int64_t e12 = eT1.shape[1];
  const int64_t e56 = eT6.shape[0];
  const int64_t e78 = eT8.shape[0];
  const int64_t t12 = Tn1.shape[2];
  const int64_t t41 = Tn1.shape[3];
  const int64_t t34 = Tn4.shape[2];
  if (t41 == 1) {
    job.PU = identity_projector(ctx, e78, t41);
    job.PL = job.PU;
    job.done = true;
    return;
  }
  tensor LT = enlarged_LT(C1, eT1, eT8, Tn1);
  tensor LB = enlarged_LB(C4, eT7, eT6, Tn4);
  const bool rsvd = (t12 != 1 && t34 != 1) && pp.Use_RSVD;
  // bra bond dimensions (equal to the ket's, or 1 in single-layer mode)
  const int64_t b12 = LT.shape[2], b41 = LT.shape[5], b34 = LB.shape[5];
  int cap = pp.CHI;
  if (rsvd)
    cap = (int)std::min<int64_t>(
        std::min<int64_t>(std::min<int64_t>(pp.CHI, e78 * t41 * b41),
                          e12 * t12 * b12),
        e56 * t34 * b34);
  if (rsvd) {
    // Mult_col / Mult_row (ctm.cpp:59-119): the RSVD branch never forms R;
    // the range finder multiplies by LT (LB X) and LB^T (LT^T Y)
    const int64_t nr = t12 * e12 * b12, nc = t34 * e56 * b34, inner = LT.size() / nr;
    const tensor LTm = tn::reshape(LT, Shape{nr, inner}), LBm = tn::reshape(LB, Shape{inner, nc});
    const int mode = r_contraction_mode(pp);
    tn::LinearOp op;
    op.m = nr;
    op.n = nc;
    op.mul = [&ctx, LTm, LBm, mode](const tensor &X) {
      tn::ContractionPrecision prec(ctx, mode);
      return tn::tensordot(LTm, tn::tensordot(LBm, X, Axes{1}, Axes{0}), Axes{1}, Axes{0});
    };
    op.mulT = [&ctx, LTm, LBm, mode](const tensor &Y) {
      tn::ContractionPrecision prec(ctx, mode);
      return tn::tensordot(LBm, tn::tensordot(LTm, Y, Axes{0}, Axes{0}), Axes{0}, Axes{0});
    };
    projectors_begin(ctx, tensor(), t12, e12, b12, t34, e56, b34, cap, LT, Axes{0, 1, 2}, LB,
                     Axes{3, 4, 5}, pp, rsvd, slot, key, job, &op);
    return;
  }
  // R = LT * LB, legs (t12, e12, tc12, t34, e56, tc34)  (ctm.cpp:273); in
  // mixed mode 1 the matrix the SVD truncates stays FP64
  tensor R;
  {
    tn::ContractionPrecision prec(ctx, r_contraction_mode(pp));
    R = tn::tensordot(LT, LB, Axes{3, 4, 5}, Axes{0, 1, 2});
  }
  projectors_begin(ctx, R, t12, e12, b12, t34, e56, b34, cap, LT,
                   Axes{0, 1, 2}, LB,
                   Axes{3, 4, 5}, pp, rsvd, slot, key, job);
}

// ctm.cpp:320-491 up to the SVD
void updown_blocks_begin(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &eT7, const tensor &eT8,
    const tensor &Tn1, const tensor &Tn2, const tensor &Tn3, const tensor &Tn4,
    const PEPS_Parameters &pp, int slot, const ProjKey *key, ProjectorJob &job) {
  tn::Context &ctx = *C1.ctx();
  const int64_t e34 = eT3.shape[1];
  const int64_t e78 = eT7.shape[1];
  const int64_t t23 = Tn2.shape[3];
  const int64_t t41 = Tn4.shape[1];
  if (t41 == 1) {
    job.PU = identity_projector(ctx, e78, t41);
    job.PL = job.PU;
    job.done = true;
    return;
  }
  tensor LT = enlarged_LT(C1, eT1, eT8, Tn1);
  // RT legs (t12, t23, e34, e12, tc12, tc23)  (ctm.cpp:364-369)
  tensor RT = tn::tensordot(
      Tn2,
      tn::tensordot(tn::tensordot(tn::tensordot(C2, eT3, Axes{1}, Axes{0}),
                                  eT2, Axes{0}, Axes{1}),
                    tn::bra(Tn2), Axes{2, 5}, Axes{2, 1}),
      Axes{1, 2, 4}, Axes{3, 1, 6});
  // RB legs (t34, t23, e56, e34, tc34, tc23)  (ctm.cpp:378-383)
  tensor RB = tn::tensordot(
      Tn3,
      tn::tensordot(tn::tensordot(tn::tensordot(C3, eT5, Axes{1}, Axes{0}),
                                  eT4, Axes{0}, Axes{1}),
                    tn::bra(Tn3), Axes{2, 5}, Axes{3, 2}),
      Axes{2, 3, 4}, Axes{3, 1, 6});
  tensor LB = enlarged_LB(C4, eT7, eT6, Tn4);
  // The single-layer network (ctm_single.cpp:267-436) takes the same blocks
  // and products with unit bra legs; its full-SVD branch truncates with a
  // loop over the first e78 singular values (ctm_single.cpp:400), its RSVD
  // branch as the double layer.
  const bool rsvd = t23 != 1 && pp.Use_RSVD;
  job.e78_loop = (ctx.single_layer() && !rsvd) ? (int)e78 : 0;
  if (rsvd) {
    // Mult_col_ud / Mult_row_ud (ctm.cpp:121-172): M = (RT LT)(RB LB)^T is
    // applied block by block, RT (LT (LB (RB X))), and neither M nor RT LT,
    // RB LB is formed; the projectors follow the same chains (ctm.cpp:400-439):
    // PL = LT (RT U), PU = LB (RB V).  Stored legs: RT (t12, t23, e34, e12,
    // tc12, tc23), RB (t34, t23, e56, e34, tc34, tc23), LT (t12, e12, tc12 |
    // t41, e78, tc41), LB (t41, e78, tc41 | t34, e56, tc34).
    const int64_t b23r = RT.shape[5], b23c = RB.shape[5];
    const int cap = (int)std::min<int64_t>(pp.CHI, e34 * t23 * b23r);
    const int mode = r_contraction_mode(pp);
    tn::LinearOp op;
    op.m = t23 * e34 * b23r;
    op.n = t23 * e34 * b23c;
    op.mul = [&ctx, LT, RT, RB, LB, mode, t23, e34, b23c](const tensor &X) {
      tn::ContractionPrecision prec(ctx, mode);
      const int64_t p = X.shape[1];
      tensor W = tn::tensordot(RB, tn::reshape(X, Shape{t23, e34, b23c, p}), Axes{1, 3, 5},
                               Axes{0, 1, 2});                            // (t34, e56, tc34, p)
      W = tn::tensordot(LB, W, Axes{3, 4, 5}, Axes{0, 1, 2});            // (t41, e78, tc41, p)
      W = tn::tensordot(LT, W, Axes{3, 4, 5}, Axes{0, 1, 2});            // (t12, e12, tc12, p)
      W = tn::tensordot(RT, W, Axes{0, 3, 4}, Axes{0, 1, 2});            // (t23, e34, tc23, p)
      return tn::reshape(W, Shape{W.size() / p, p});
    };
    op.mulT = [&ctx, LT, RT, RB, LB, mode, t23, e34, b23r](const tensor &Y) {
      tn::ContractionPrecision prec(ctx, mode);
      const int64_t p = Y.shape[1];
      tensor W = tn::tensordot(RT, tn::reshape(Y, Shape{t23, e34, b23r, p}), Axes{1, 2, 5},
                               Axes{0, 1, 2});                            // (t12, e12, tc12, p)
      W = tn::tensordot(LT, W, Axes{0, 1, 2}, Axes{0, 1, 2});            // (t41, e78, tc41, p)
      W = tn::tensordot(LB, W, Axes{0, 1, 2}, Axes{0, 1, 2});            // (t34, e56, tc34, p)
      W = tn::tensordot(RB, W, Axes{0, 2, 4}, Axes{0, 1, 2});            // (t23, e34, tc23, p)
      return tn::reshape(W, Shape{W.size() / p, p});
    };
    projectors_begin(ctx, tensor(), t23, e34, b23r, t23, e34, b23c, cap, tensor(), Axes{}, tensor(),
                     Axes{}, pp, rsvd, slot, key, job, &op);
    job.finish_a = [LT, RT](const tensor &U) {
      return tn::tensordot_t(LT, tn::tensordot(RT, U, Axes{1, 2, 5}, Axes{0, 1, 2}), Axes{0, 1, 2},
                             Axes{0, 1, 2}, Axes{1, 0, 2, 3});
    };
    job.finish_b = [LB, RB](const tensor &V) {
      return tn::tensordot_t(LB, tn::tensordot(RB, V, Axes{1, 3, 5}, Axes{0, 1, 2}), Axes{3, 4, 5},
                             Axes{0, 1, 2}, Axes{1, 0, 2, 3});
    };
    return;
  }
  // R1 = RT*LT, R2 = RB*LB, legs (t23, e34, tc23, t41, e78, tc41)
  tensor R1 = tn::tensordot(RT, LT, Axes{0, 3, 4}, Axes{0, 1, 2});
  tensor R2 = tn::tensordot(RB, LB, Axes{0, 2, 4}, Axes{3, 4, 5});
  const int64_t b23 = R1.shape[2];
  int cap = pp.CHI;
  if (rsvd) cap = (int)std::min<int64_t>(pp.CHI, e34 * t23 * b23);
  tensor M;
  {
    tn::ContractionPrecision prec(ctx, r_contraction_mode(pp));
    M = tn::tensordot(R1, R2, Axes{3, 4, 5}, Axes{3, 4, 5});
  }
  projectors_begin(ctx, M, t23, e34, b23, t23, e34, R2.shape[2], cap, R1,
                   Axes{0, 1, 2}, R2, Axes{0, 1, 2}, pp, rsvd, slot, key, job);
}

}  // namespace

// ctm.cpp:174-318
void Calc_projector_left_block(const tensor &C1, const tensor &C4,
                               const tensor &eT1, const tensor &eT6,
                               const tensor &eT7, const tensor &eT8,
                               const tensor &Tn1, const tensor &Tn4,
                               const PEPS_Parameters &pp, tensor &PU,
                               tensor &PL) {
  ProjectorJob job;
  left_block_begin(C1, C4, eT1, eT6, eT7, eT8, Tn1, Tn4, pp, kSyncSlot, nullptr,
                   job);
  projectors_finish(*C1.ctx(), job, pp);
  PU = job.PU;
  PL = job.PL;
}

// ctm.cpp:320-491
void Calc_projector_updown_blocks(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &eT7, const tensor &eT8,
    const tensor &Tn1, const tensor &Tn2, const tensor &Tn3, const tensor &Tn4,
    const PEPS_Parameters &pp, tensor &PU, tensor &PL) {
  ProjectorJob job;
  updown_blocks_begin(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, eT7, eT8,
                      Tn1, Tn2, Tn3, Tn4, pp, kSyncSlot, nullptr, job);
  projectors_finish(*C1.ctx(), job, pp);
  PU = job.PU;
  PL = job.PL;
}

// ctm.cpp:493-518
void Calc_Next_CTM(const tensor &C1, const tensor &C4, const tensor &eT1,
                   const tensor &eT6, const tensor &PU, const tensor &PL,
                   tensor &C1_out, tensor &C4_out) {
  C1_out = tn::tensordot(PU, tn::tensordot(C1, eT1, Axes{1}, Axes{0}),
                         Axes{0, 1, 2}, Axes{0, 2, 3});
  C4_out = tn::tensordot(tn::tensordot(eT6, C4, Axes{1}, Axes{0}), PL,
                         Axes{3, 1, 2}, Axes{0, 1, 2});
  tn::normalize_max_abs(C1_out);
  tn::normalize_max_abs(C4_out);
}

// ctm.cpp:520-552: ((Tn1*(Tn1c*(eT8*PL)))*PU), then transpose (3,2,0,1)
// fused into the last contraction's store.
void Calc_Next_eT(const tensor &eT8, const tensor &Tn1, const tensor &PU,
                  const tensor &PL, tensor &eT_out) {
  if (fused_site_ok(Tn1)) {
    // (a, c, k8, t, b8, tc): eT8's ket / bra legs meet Tn1's l, PL's its u
    const tensor X = tn::tensordot_t(eT8, PL, Axes{1}, Axes{0}, Axes{0, 5, 1, 3, 2, 4});
    const tensor Z = fused_site(X, Tn1, 0, 1, true);  // (a, c, r, d, r', d')
    eT_out = tn::tensordot_t(Z, PU, Axes{3, 5, 0}, Axes{1, 2, 0}, Axes{3, 0, 1, 2});
  } else {
    const tensor X = tn::tensordot(eT8, PL, Axes{1}, Axes{0});  // (a, k8, b8, t, tc, c)
    const tensor Y = tn::tensordot(tn::bra(Tn1), X, Axes{0, 1}, Axes{2, 4});
    const tensor Z = tn::tensordot(Tn1, Y, Axes{0, 1, 4}, Axes{4, 5, 2});  // (r, d, r', d', a, c)
    eT_out = tn::tensordot_t(Z, PU, Axes{1, 3, 4}, Axes{1, 2, 0}, Axes{3, 2, 0, 1});
  }
  tn::normalize_max_abs(eT_out);
}

namespace {

// One directional move split at its projector synchronisation: begin(row,
// slot, job) issues the projector of a row; once all are done, snapshot()
// keeps the arrays the renormalisation reads, and update(row, PUs, PLs)
// issues one row's renormalisation (Calc_Next_CTM and two Calc_Next_eT).
struct MoveSpec {
  int L = 0;
  std::function<void(int, int, ProjectorJob &)> begin;
  std::function<void()> snapshot;
  std::function<void(int, const std::vector<tensor> &,
                     const std::vector<tensor> &)>
      update;
};

// Runs one move, or two moves that read and write disjoint arrays, with all
// their projector chains overlapped on the lanes of the forked context.
// With the cut on the device the singular values of projector q go to pinned
// slot `slot` + q (slot advances by the number of projectors) and are read
// at the caller's next synchronisation (collect_svd_stats).  slot_end bounds
// the slots: the uncaptured range (kEagerSlots; when it is used up the
// slots are read first and counting restarts at 0) or a captured sweep's
// window (which holds all its projectors).
void run_moves(tn::Context &ctx, const PEPS_Parameters &pp,
               const std::vector<MoveSpec *> &moves, int &slot,
               int slot_end = tn::Context::kEagerSlots) {
  tn::ContractionPrecision prec(ctx, move_contraction_mode(pp));
  const bool dev = pp.projector_cut_on_device;
  int total = 0;
  for (auto *m : moves) total += m->L;
  TN_REQUIRE(total <= tn::Context::kEagerSlots, "CTM move: too many projectors in one move");
  if (dev && slot + total > slot_end) {
    TN_REQUIRE(slot_end == tn::Context::kEagerSlots,
               "CTM move: the captured sweep's pinned slot window is too small");
    // the pinned slots are used up: read them first
    ctx.sync();
    ctx.collect_svd_stats();
    slot = 0;
  }
  const int nl = std::min(total, ctx.max_lanes());
  const bool fork = nl > 1 && !ctx.forked();
  std::vector<std::vector<ProjectorJob>> jobs(moves.size());
  std::vector<std::vector<tensor>> PUs(moves.size()), PLs(moves.size());
  {
    tn::ForkScope lanes(ctx, nl, fork);
    int q = 0;
    for (size_t m = 0; m < moves.size(); ++m) {
      jobs[m].resize(moves[m]->L);
      for (int i = 0; i < moves[m]->L; ++i, ++q) {
        if (fork) ctx.use_lane(q % nl);
        jobs[m][i].device_cut = dev;
        moves[m]->begin(i, dev ? slot + q : (fork ? q : kSyncSlot), jobs[m][i]);
        if (!fork) projectors_finish(ctx, jobs[m][i], pp);
      }
    }
    if (dev) slot += total;
    q = 0;
    for (size_t m = 0; m < moves.size(); ++m) {
      PUs[m].resize(moves[m]->L);
      PLs[m].resize(moves[m]->L);
      for (int i = 0; i < moves[m]->L; ++i, ++q) {
        if (fork) ctx.use_lane(q % nl);
        projectors_finish(ctx, jobs[m][i], pp);
        PUs[m][i] = jobs[m][i].PU;
        PLs[m][i] = jobs[m][i].PL;
      }
    }
    // a row's renormalisation uses the projectors of the neighbouring rows
    lanes.end();
  }
  for (auto *m : moves) m->snapshot();
  tn::ForkScope lanes(ctx, nl, fork);
  int q = 0;
  for (size_t m = 0; m < moves.size(); ++m)
    for (int i = 0; i < moves[m]->L; ++i, ++q) {
      if (fork) ctx.use_lane(q % nl);
      moves[m]->update(i, PUs[m], PLs[m]);
    }
  lanes.end();
  ctx.move_issued();
}

using Arr = std::vector<tensor>;

// ctm.cpp:562-616
MoveSpec left_spec(Arr &C1, const Arr &C2, const Arr &C3, Arr &C4,
                   const Arr &eTt, const Arr &eTr, const Arr &eTb, Arr &eTl,
                   const Arr &Tn, int ix, const PEPS_Parameters &pp,
                   const SquareLattice &lat) {
  MoveSpec m;
  const int L = m.L = lat.LY_noskew;
  auto bak = std::make_shared<std::array<Arr, 3>>();
  m.begin = [&, ix](int iy, int slot, ProjectorJob &job) {
    const ProjKey key{environment_id(C1), 0, ix, iy};
    const int i = lat.index(ix, iy);
    const int j = lat.right(i);
    const int k = lat.bottom(j);
    const int l = lat.left(k);
    if (pp.CTM_Projector_corner)
      left_block_begin(C1[i], C4[l], eTt[i], eTb[l], eTl[l], eTl[i], Tn[i],
                       Tn[l], pp, slot, &key, job);
    else
      updown_blocks_begin(C1[i], C2[j], C3[k], C4[l], eTt[i], eTt[j], eTr[j],
                          eTr[k], eTb[k], eTb[l], eTl[l], eTl[i], Tn[i],
                          Tn[j], Tn[k], Tn[l], pp, slot, &key, job);
  };
  m.snapshot = [&, bak] { *bak = {C1, C4, eTl}; };
  m.update = [&, ix, L, bak](int iy, const Arr &PUs, const Arr &PLs) {
    const Arr &C1_bak = (*bak)[0], &C4_bak = (*bak)[1], &eTl_bak = (*bak)[2];
    const int i = lat.index(ix, iy);
    const int j = lat.right(i);
    const int k = lat.bottom(j);
    const int l = lat.left(k);
    const int iy_up = (iy + 1) % L;
    const int iy_down = (iy - 1 + L) % L;
    Calc_Next_CTM(C1_bak[i], C4_bak[l], eTt[i], eTb[l], PUs[iy_up],
                  PLs[iy_down], C1[j], C4[k]);
    Calc_Next_eT(eTl_bak[i], Tn[i], PUs[iy], PLs[iy_up], eTl[j]);
    Calc_Next_eT(eTl_bak[l], Tn[l], PUs[iy_down], PLs[iy], eTl[k]);
  };
  return m;
}

// ctm.cpp:618-681
MoveSpec right_spec(const Arr &C1, Arr &C2, Arr &C3, const Arr &C4,
                    const Arr &eTt, Arr &eTr, const Arr &eTb, const Arr &eTl,
                    const Arr &Tn, int ix, const PEPS_Parameters &pp,
                    const SquareLattice &lat) {
  MoveSpec m;
  const int L = m.L = lat.LY_noskew;
  const Axes rot{2, 3, 0, 1, 4};
  auto bak = std::make_shared<std::array<Arr, 3>>();
  m.begin = [&, ix, rot](int iy, int slot, ProjectorJob &job) {
    const ProjKey key{environment_id(C1), 1, ix, iy};
    const int k = lat.index(ix, iy);
    const int l = lat.left(k);
    const int i = lat.top(l);
    const int j = lat.right(i);
    if (pp.CTM_Projector_corner)
      left_block_begin(C3[k], C2[j], eTb[k], eTt[j], eTr[j], eTr[k],
                       tn::transpose(Tn[k], rot), tn::transpose(Tn[j], rot),
                       pp, slot, &key, job);
    else
      updown_blocks_begin(
          C3[k], C4[l], C1[i], C2[j], eTb[k], eTb[l], eTl[l], eTl[i], eTt[i],
          eTt[j], eTr[j], eTr[k], tn::transpose(Tn[k], rot),
          tn::transpose(Tn[l], rot), tn::transpose(Tn[i], rot),
          tn::transpose(Tn[j], rot), pp, slot, &key, job);
  };
  m.snapshot = [&, bak] { *bak = {C2, C3, eTr}; };
  m.update = [&, ix, L, rot, bak](int iy, const Arr &PUs, const Arr &PLs) {
    const Arr &C2_bak = (*bak)[0], &C3_bak = (*bak)[1], &eTr_bak = (*bak)[2];
    const int k = lat.index(ix, iy);
    const int l = lat.left(k);
    const int i = lat.top(l);
    const int j = lat.right(i);
    const int iy_up = (iy + 1) % L;
    const int iy_down = (iy - 1 + L) % L;
    Calc_Next_CTM(C3_bak[k], C2_bak[j], eTb[k], eTt[j], PUs[iy_down],
                  PLs[iy_up], C3[l], C2[i]);
    Calc_Next_eT(eTr_bak[k], tn::transpose(Tn[k], rot), PUs[iy],
                 PLs[iy_down], eTr[l]);
    Calc_Next_eT(eTr_bak[j], tn::transpose(Tn[j], rot), PUs[iy_up], PLs[iy],
                 eTr[i]);
  };
  return m;
}

// ctm.cpp:683-746
MoveSpec top_spec(Arr &C1, Arr &C2, const Arr &C3, const Arr &C4, Arr &eTt,
                  const Arr &eTr, const Arr &eTb, const Arr &eTl,
                  const Arr &Tn, int iy, const PEPS_Parameters &pp,
                  const SquareLattice &lat) {
  MoveSpec m;
  const int L = m.L = lat.LX_noskew;
  const Axes rot{1, 2, 3, 0, 4};
  auto bak = std::make_shared<std::array<Arr, 3>>();
  m.begin = [&, iy, rot](int ix, int slot, ProjectorJob &job) {
    const ProjKey key{environment_id(C1), 2, iy, ix};
    const int j = lat.index(ix, iy);
    const int k = lat.bottom(j);
    const int l = lat.left(k);
    const int i = lat.top(l);
    if (pp.CTM_Projector_corner)
      left_block_begin(C2[j], C1[i], eTr[j], eTl[i], eTt[i], eTt[j],
                       tn::transpose(Tn[j], rot), tn::transpose(Tn[i], rot),
                       pp, slot, &key, job);
    else
      updown_blocks_begin(
          C2[j], C3[k], C4[l], C1[i], eTr[j], eTr[k], eTb[k], eTb[l], eTl[l],
          eTl[i], eTt[i], eTt[j], tn::transpose(Tn[j], rot),
          tn::transpose(Tn[k], rot), tn::transpose(Tn[l], rot),
          tn::transpose(Tn[i], rot), pp, slot, &key, job);
  };
  m.snapshot = [&, bak] { *bak = {C1, C2, eTt}; };
  m.update = [&, iy, L, rot, bak](int ix, const Arr &PUs, const Arr &PLs) {
    const Arr &C1_bak = (*bak)[0], &C2_bak = (*bak)[1], &eTt_bak = (*bak)[2];
    const int j = lat.index(ix, iy);
    const int k = lat.bottom(j);
    const int l = lat.left(k);
    const int i = lat.top(l);
    const int ix_right = (ix + 1) % L;
    const int ix_left = (ix - 1 + L) % L;
    Calc_Next_CTM(C2_bak[j], C1_bak[i], eTr[j], eTl[i], PUs[ix_right],
                  PLs[ix_left], C2[k], C1[l]);
    Calc_Next_eT(eTt_bak[j], tn::transpose(Tn[j], rot), PUs[ix],
                 PLs[ix_right], eTt[k]);
    Calc_Next_eT(eTt_bak[i], tn::transpose(Tn[i], rot), PUs[ix_left], PLs[ix],
                 eTt[l]);
  };
  return m;
}

// ctm.cpp:748-813
MoveSpec bottom_spec(const Arr &C1, const Arr &C2, Arr &C3, Arr &C4,
                     const Arr &eTt, const Arr &eTr, Arr &eTb, const Arr &eTl,
                     const Arr &Tn, int iy, const PEPS_Parameters &pp,
                     const SquareLattice &lat) {
  MoveSpec m;
  const int L = m.L = lat.LX_noskew;
  const Axes rot{3, 0, 1, 2, 4};
  auto bak = std::make_shared<std::array<Arr, 3>>();
  m.begin = [&, iy, rot](int ix, int slot, ProjectorJob &job) {
    const ProjKey key{environment_id(C1), 3, iy, ix};
    const int l = lat.index(ix, iy);
    const int i = lat.top(l);
    const int j = lat.right(i);
    const int k = lat.bottom(j);
    if (pp.CTM_Projector_corner)
      left_block_begin(C4[l], C3[k], eTl[l], eTr[k], eTb[k], eTb[l],
                       tn::transpose(Tn[l], rot), tn::transpose(Tn[k], rot),
                       pp, slot, &key, job);
    else
      updown_blocks_begin(
          C4[l], C1[i], C2[j], C3[k], eTl[l], eTl[i], eTt[i], eTt[j], eTr[j],
          eTr[k], eTb[k], eTb[l], tn::transpose(Tn[l], rot),
          tn::transpose(Tn[i], rot), tn::transpose(Tn[j], rot),
          tn::transpose(Tn[k], rot), pp, slot, &key, job);
  };
  m.snapshot = [&, bak] { *bak = {C3, C4, eTb}; };
  m.update = [&, iy, L, rot, bak](int ix, const Arr &PUs, const Arr &PLs) {
    const Arr &C3_bak = (*bak)[0], &C4_bak = (*bak)[1], &eTb_bak = (*bak)[2];
    const int l = lat.index(ix, iy);
    const int i = lat.top(l);
    const int j = lat.right(i);
    const int k = lat.bottom(j);
    const int ix_right = (ix + 1) % L;
    const int ix_left = (ix - 1 + L) % L;
    Calc_Next_CTM(C4_bak[l], C3_bak[k], eTl[l], eTr[k], PUs[ix_left],
                  PLs[ix_right], C4[i], C3[j]);
    Calc_Next_eT(eTb_bak[l], tn::transpose(Tn[l], rot), PUs[ix], PLs[ix_left],
                 eTb[i]);
    Calc_Next_eT(eTb_bak[k], tn::transpose(Tn[k], rot), PUs[ix_right],
                 PLs[ix], eTb[j]);
  };
  return m;
}

}  // namespace

// One move on its own (the public move functions): with the cut on the
// device the move ends with a synchronisation that reads its projectors'
// statistics.
static void run_single_move(tn::Context &ctx, const PEPS_Parameters &pp, MoveSpec &m) {
  int slot = ctx.eager_slot;
  run_moves(ctx, pp, {&m}, slot);
  ctx.eager_slot = slot;
  if (pp.projector_cut_on_device) {
    ctx.sync();
    ctx.collect_svd_stats();
  }
}

// ctm.cpp:562-616
void Left_move(std::vector<tensor> &C1, const std::vector<tensor> &C2,
               const std::vector<tensor> &C3, std::vector<tensor> &C4,
               const std::vector<tensor> &eTt, const std::vector<tensor> &eTr,
               const std::vector<tensor> &eTb, std::vector<tensor> &eTl,
               const std::vector<tensor> &Tn, int ix,
               const PEPS_Parameters &pp, const SquareLattice &lat) {
  MoveSpec m = left_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, ix, pp, lat);
  run_single_move(*Tn[0].ctx(), pp, m);
}

// ctm.cpp:618-681
void Right_move(const std::vector<tensor> &C1, std::vector<tensor> &C2,
                std::vector<tensor> &C3, const std::vector<tensor> &C4,
                const std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                const std::vector<tensor> &eTb, const std::vector<tensor> &eTl,
                const std::vector<tensor> &Tn, int ix,
                const PEPS_Parameters &pp, const SquareLattice &lat) {
  MoveSpec m = right_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, ix, pp, lat);
  run_single_move(*Tn[0].ctx(), pp, m);
}

// ctm.cpp:683-746
void Top_move(std::vector<tensor> &C1, std::vector<tensor> &C2,
              const std::vector<tensor> &C3, const std::vector<tensor> &C4,
              std::vector<tensor> &eTt, const std::vector<tensor> &eTr,
              const std::vector<tensor> &eTb, const std::vector<tensor> &eTl,
              const std::vector<tensor> &Tn, int iy, const PEPS_Parameters &pp,
              const SquareLattice &lat) {
  MoveSpec m = top_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, iy, pp, lat);
  run_single_move(*Tn[0].ctx(), pp, m);
}

// ctm.cpp:748-813
void Bottom_move(const std::vector<tensor> &C1, const std::vector<tensor> &C2,
                 std::vector<tensor> &C3, std::vector<tensor> &C4,
                 const std::vector<tensor> &eTt,
                 const std::vector<tensor> &eTr, std::vector<tensor> &eTb,
                 const std::vector<tensor> &eTl, const std::vector<tensor> &Tn,
                 int iy, const PEPS_Parameters &pp, const SquareLattice &lat) {
  MoveSpec m = bottom_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, iy, pp, lat);
  run_single_move(*Tn[0].ctx(), pp, m);
}

namespace {

// L2-normalised descending spectra of a list of matrices, batched by shape.
std::vector<std::vector<double>> spectra(const std::vector<const tensor *> &Cs) {
  std::vector<std::vector<double>> out(Cs.size());
  if (Cs.empty()) return out;
  tn::Context &ctx = *Cs[0]->ctx();
  std::map<Shape, std::vector<int>> groups;
  for (size_t i = 0; i < Cs.size(); ++i) groups[Cs[i]->shape].push_back((int)i);
  for (auto &g : groups) {
    std::vector<tensor> mats;
    for (int i : g.second) mats.push_back(*Cs[i]);
    bool big = g.first[0] > tn::kern::jacobi_max_dim() ||
               g.first[1] > tn::kern::jacobi_max_dim();
    TN_REQUIRE(!big, "convergence check: corner matrices above 256 x 256 are "
                     "not supported");
    tensor sv = tn::svd_values_batched(ctx, mats);
    const int64_t k = sv.shape[1];
    std::vector<double> h((size_t)sv.size());
    tn::to_host(sv, h.data());
    for (size_t q = 0; q < g.second.size(); ++q) {
      std::vector<double> lam(h.begin() + q * k, h.begin() + (q + 1) * k);
      double norm = 0.0;
      for (double v : lam) norm += v * v;
      norm = std::sqrt(norm);
      for (double &v : lam) v /= norm;
      out[g.second[q]] = lam;
    }
  }
  return out;
}

}  // namespace

std::vector<double> normalized_spectrum(const tensor &C) {
  return spectra({&C})[0];
}

namespace {
// reshape(...) then extend/slice to (CHI, CHI[, d, d]) (ctm.cpp:998-1012)
tensor fit_chi(const tensor &T, int64_t chi) {
  tensor r = T;
  const int64_t d1 = T.shape[0], d2 = T.shape[1];
  if (d1 > chi) r = tn::slice(r, 0, 0, chi);
  if (d2 > chi) r = tn::slice(r, 1, 0, chi);
  if (r.shape[0] < chi || r.shape[1] < chi) {
    Shape s = r.shape;
    s[0] = chi;
    s[1] = chi;
    r = tn::extend(r, s);
  }
  return r;
}
}  // namespace

// ctm.cpp:984-1193
// The bra copy of T in the initial corner and edge contractions: conj(T),
// or in single-layer mode a tensor of ones over the contracted legs (unit
// over the others), which sums T over those legs (the uniform vectors of
// ctm_single.cpp:751-995).
tensor init_bra(const tensor &T, const Axes &contracted) {
  if (!T.ctx()->single_layer()) return tn::conj(T);
  Shape s(5, 1);
  for (int k : contracted) s[k] = T.shape[k];
  tensor B = tn::empty(*T.ctx(), s);
  tn::kern::fill(B.data(), B.size(), 1.0, T.ctx()->stream());
  return B;
}

void Initialize_CTM(std::vector<tensor> &C1, std::vector<tensor> &C2,
                    std::vector<tensor> &C3, std::vector<tensor> &C4,
                    std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                    std::vector<tensor> &eTb, std::vector<tensor> &eTl,
                    const std::vector<tensor> &Tn, const PEPS_Parameters &pp,
                    const SquareLattice &lat) {
  const int N = lat.N_UNIT;
  const int64_t chi = pp.CHI;
  // the single-layer initial environment keeps the site tensors' bond
  // dimensions (ctm_single.cpp:751-995 has no extend/slice to chi)
  const bool single = Tn[0].ctx()->single_layer();
  auto fit = [&](const tensor &x) { return single ? x : fit_chi(x, chi); };
  C1.resize(N);
  C2.resize(N);
  C3.resize(N);
  C4.resize(N);
  eTt.resize(N);
  eTr.resize(N);
  eTb.resize(N);
  eTl.resize(N);
  for (int i = 0; i < N; ++i) {
    int num = lat.top(lat.left(i));
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{0, 1, 4});
      tensor c = tn::tensordot_t(T, B, Axes{0, 1, 4}, Axes{0, 1, 4},
                                 Axes{1, 3, 0, 2});
      C1[i] = fit(tn::reshape(c, Shape{T.shape[3] * B.shape[3],
                                       T.shape[2] * B.shape[2]}));
    }
    num = lat.top(lat.right(i));
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{1, 2, 4});
      tensor c = tn::tensordot_t(T, B, Axes{1, 2, 4}, Axes{1, 2, 4},
                                 Axes{0, 2, 1, 3});
      C2[i] = fit(tn::reshape(c, Shape{T.shape[0] * B.shape[0],
                                       T.shape[3] * B.shape[3]}));
    }
    num = lat.bottom(lat.right(i));
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{2, 3, 4});
      tensor c = tn::tensordot_t(T, B, Axes{2, 3, 4}, Axes{2, 3, 4},
                                 Axes{1, 3, 0, 2});
      C3[i] = fit(tn::reshape(c, Shape{T.shape[1] * B.shape[1],
                                       T.shape[0] * B.shape[0]}));
    }
    num = lat.bottom(lat.left(i));
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{0, 3, 4});
      tensor c = tn::tensordot_t(T, B, Axes{0, 3, 4}, Axes{0, 3, 4},
                                 Axes{1, 3, 0, 2});
      C4[i] = fit(tn::reshape(c, Shape{T.shape[2] * B.shape[2],
                                       T.shape[1] * B.shape[1]}));
    }
    num = lat.top(i);
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{1, 4});
      tensor e = tn::tensordot_t(T, B, Axes{1, 4}, Axes{1, 4},
                                 Axes{0, 3, 1, 4, 2, 5});
      eTt[i] = fit(tn::reshape(e, Shape{T.shape[0] * B.shape[0],
                                        T.shape[2] * B.shape[2],
                                        T.shape[3], B.shape[3]}));
    }
    num = lat.right(i);
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{2, 4});
      tensor e = tn::tensordot_t(T, B, Axes{2, 4}, Axes{2, 4},
                                 Axes{1, 4, 2, 5, 0, 3});
      eTr[i] = fit(tn::reshape(e, Shape{T.shape[1] * B.shape[1],
                                        T.shape[3] * B.shape[3],
                                        T.shape[0], B.shape[0]}));
    }
    num = lat.bottom(i);
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{3, 4});
      tensor e = tn::tensordot_t(T, B, Axes{3, 4}, Axes{3, 4},
                                 Axes{2, 5, 0, 3, 1, 4});
      eTb[i] = fit(tn::reshape(e, Shape{T.shape[2] * B.shape[2],
                                        T.shape[0] * B.shape[0],
                                        T.shape[1], B.shape[1]}));
    }
    num = lat.left(i);
    {
      const auto &T = Tn[num];
      const tensor B = init_bra(T, Axes{0, 4});
      tensor e = tn::tensordot_t(T, B, Axes{0, 4}, Axes{0, 4},
                                 Axes{2, 5, 0, 3, 1, 4});
      eTl[i] = fit(tn::reshape(e, Shape{T.shape[3] * B.shape[3],
                                        T.shape[1] * B.shape[1],
                                        T.shape[2], B.shape[2]}));
    }
  }
}

namespace {

// The moves of one sweep with corner projectors.  The left moves read and
// write only C1, C4, eTl (and read eTt, eTb, Tn), the right moves only C2,
// C3, eTr: the two chains of ctm.cpp:1205-1214 are independent and run side
// by side, step s of one with step s of the other; likewise the top (C1, C2,
// eTt) and bottom (C3, C4, eTb) chains of ctm.cpp:1215-1224.  Each chain
// keeps the reference's order, so the result is the same.
// With window_first >= 0 the projectors' pinned slots are the window
// starting there (a captured sweep's own, sweep_projectors(lat) long),
// else the context's uncaptured slots.
int sweep_projectors(const SquareLattice &lat) {
  return 2 * lat.LX_noskew * lat.LY_noskew + 2 * lat.LY_noskew * lat.LX_noskew;
}

void corner_sweep(std::vector<tensor> *A[8], const std::vector<tensor> &Tn,
                  const PEPS_Parameters &ps, const SquareLattice &lat,
                  int window_first = -1) {
  auto &C1 = *A[0], &C2 = *A[1], &C3 = *A[2], &C4 = *A[3];
  auto &eTt = *A[4], &eTr = *A[5], &eTb = *A[6], &eTl = *A[7];
  tn::Context &ctx = *Tn[0].ctx();
  const int LX = lat.LX_noskew, LY = lat.LY_noskew;
  const bool window = window_first >= 0;
  int slot = window ? window_first : ctx.eager_slot;
  const int slot_end = window ? window_first + sweep_projectors(lat) : tn::Context::kEagerSlots;
  for (int s = 0; s < LX; ++s) {
    MoveSpec a = left_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, s, ps, lat);
    MoveSpec b = right_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn,
                            (-s + 1 + LX) % LX, ps, lat);
    run_moves(ctx, ps, {&a, &b}, slot, slot_end);
  }
  for (int s = 0; s < LY; ++s) {
    MoveSpec a = top_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn,
                          (-s + 1 + LY) % LY, ps, lat);
    MoveSpec b = bottom_spec(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, s, ps, lat);
    run_moves(ctx, ps, {&a, &b}, slot, slot_end);
  }
  if (!window) ctx.eager_slot = slot;
}

// What a captured sweep of an environment bakes in: the buffers and shapes
// of every tensor it reads and writes, the warm-start bases with the
// range-finder decision each basis's Ritz-value range implies (svd_top_begin
// skips the orthonormalisation between the A^T and A products below 1e7),
// and the parameters that shape the kernels.
std::vector<uintptr_t> sweep_signature(tn::Context &ctx, int64_t env,
                                       std::vector<tensor> *A[8],
                                       const std::vector<tensor> &Tn,
                                       const PEPS_Parameters &pp) {
  std::vector<uintptr_t> sig;
  auto add = [&](const std::vector<tensor> &arr) {
    for (const auto &t : arr) {
      sig.push_back((uintptr_t)t.data());
      for (auto d : t.shape) sig.push_back((uintptr_t)d);
      sig.push_back(~(uintptr_t)0);
    }
  };
  for (int k = 0; k < 8; ++k) add(*A[k]);
  add(Tn);
  for (const auto &kv : ctx.svd_start) {
    if (kv.first[0] != env) continue;
    for (auto x : kv.first) sig.push_back((uintptr_t)x);
    sig.push_back((uintptr_t)kv.second.V.data());
    sig.push_back((uintptr_t)kv.second.s.data());
    sig.push_back(kv.second.range > 0.0 && kv.second.range < 1e7);
  }
  uint64_t cut;
  std::memcpy(&cut, &pp.Inverse_projector_cut, sizeof(cut));
  for (uintptr_t x : {(uintptr_t)pp.CHI, (uintptr_t)cut, (uintptr_t)pp.svd_power_iters,
                      (uintptr_t)pp.rsvd_power_iters, (uintptr_t)pp.svd_warm_start,
                      (uintptr_t)pp.svd_warm_power_iters, (uintptr_t)pp.contraction_fp32,
                      (uintptr_t)pp.Use_RSVD, (uintptr_t)ctx.max_lanes()})
    sig.push_back(x);
  return sig;
}

// A sweep can be captured (or replayed) when every projector already has its
// warm-start basis (or warm starts are off) and the environment exclusively
// owns its tensors: a replay writes its results into those buffers, which no
// other handle may see change.
bool sweep_graph_ready(tn::Context &ctx, int64_t env, std::vector<tensor> *A[8],
                       const SquareLattice &lat, const PEPS_Parameters &pp) {
  for (int k = 0; k < 8; ++k)
    for (const auto &t : *A[k])
      if (!t.defined() || t.st.use_count() != 1) return false;
  if (!pp.svd_warm_start) return true;
  int n = 0;
  for (const auto &kv : ctx.svd_start) n += kv.first[0] == env;
  return n == 4 * lat.LX_noskew * lat.LY_noskew;
}

// Captures one sweep of the environment into a graph: the sweep is issued
// into the capture with every allocation from a private arena, then its
// results (the eight environment arrays and the warm-start bases) are
// copied into the buffers the sweep started from, so each replay reads and
// writes the same addresses.  Returns false (nothing captured, nothing run)
// when the capture fails; the arrays are restored in any case.
bool capture_sweep(tn::Context &ctx, int64_t env, std::vector<tensor> *A[8],
                   const std::vector<tensor> &Tn, const PEPS_Parameters &ps,
                   const SquareLattice &lat) {
  tn::Context::SweepGraph g;
  g.slot_count = sweep_projectors(lat);
  g.slot_first = ctx.slot_window(g.slot_count);
  if (g.slot_first < 0) return false;  // no pinned slots left: stay uncaptured
  g.arena = ctx.allocator().arena_create();
  std::vector<tensor> in[8];
  for (int k = 0; k < 8; ++k) in[k] = *A[k];
  std::map<std::array<int64_t, 5>, tn::Context::SvdStart> starts;
  for (const auto &kv : ctx.svd_start)
    if (kv.first[0] == env) starts.insert(kv);
  const double flops0 = ctx.gemm_flops;
  const int64_t calls0 = ctx.gemm_calls, svd0 = ctx.svd_calls;
  const size_t pend0 = ctx.pending_svd.size();
  bool ok = true;
  ctx.allocator().arena_activate(g.arena);
  try {
    ctx.begin_capture();
  } catch (...) {
    ctx.allocator().arena_activate(-1);
    ctx.release_sweep_graph(g);
    return false;
  }
  try {
    corner_sweep(A, Tn, ps, lat, g.slot_first);
    hipStream_t s = ctx.stream();
    auto copy_back = [&](const tensor &dst, const tensor &src) {
      if (dst.st == src.st) return;
      TN_REQUIRE(dst.size() == src.size(), "captured sweep: a tensor changed its size");
      TN_HIP_CHECK(hipMemcpyAsync(dst.data(), src.data(), (size_t)dst.size() * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
    };
    for (int k = 0; k < 8; ++k)
      for (size_t i = 0; i < in[k].size(); ++i) copy_back(in[k][i], (*A[k])[i]);
    for (const auto &kv : starts) {
      const auto &now = ctx.svd_start.at(kv.first);
      copy_back(kv.second.V, now.V);
      copy_back(kv.second.s, now.s);
    }
  } catch (...) {
    ok = false;
  }
  hipGraph_t graph = ctx.end_capture(g.blocks);
  ctx.allocator().arena_activate(-1);
  for (int k = 0; k < 8; ++k) *A[k] = in[k];
  for (const auto &kv : starts) ctx.svd_start[kv.first] = kv.second;
  std::vector<tn::Context::PendingSvdStat> pend(ctx.pending_svd.begin() + pend0,
                                                ctx.pending_svd.end());
  ctx.pending_svd.resize(pend0);
  g.gemm_flops = ctx.gemm_flops - flops0;
  g.gemm_calls = ctx.gemm_calls - calls0;
  g.svd_calls = ctx.svd_calls - svd0;
  ctx.gemm_flops = flops0;
  ctx.gemm_calls = calls0;
  ctx.svd_calls = svd0;
  if (ok && graph) {
    ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (!ok) (void)hipGetLastError();
  }
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) {
    ctx.release_sweep_graph(g);
    return false;
  }
  g.pending = std::move(pend);
  g.signature = sweep_signature(ctx, env, A, Tn, ps);
  ctx.sweep_graphs[env] = std::move(g);
  ctx.graph_captures++;
  return true;
}

}  // namespace

// Per-corner convergence of one environment from the spectra of its new and
// previous corners (lam[8 i + 2 c] and lam[8 i + 2 c + 1] for site i, corner
// c): the reference's order and early exit (ctm.cpp:815-969).
static bool convergence_from_spectra(const std::vector<std::vector<double>> &lam,
                                     size_t first, int N, double eps,
                                     double &sig_max, double *sig_all) {
  std::vector<double> sigs(4 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    for (int c = 0; c < 4; ++c) {
      const auto &ln = lam[first + 8 * i + 2 * c];
      const auto &lo = lam[first + 8 * i + 2 * c + 1];
      double sig = 0.0;
      for (size_t j = 0; j < ln.size(); ++j) {
        // the reference reads lam_old[j] for j < lam_new.size(); a missing
        // entry (shorter old spectrum) counts as zero here.
        const double o = j < lo.size() ? lo[j] : 0.0;
        sig += (ln[j] - o) * (ln[j] - o);
      }
      sigs[4 * i + c] = std::sqrt(sig);
    }
  }
  if (sig_all) *sig_all = *std::max_element(sigs.begin(), sigs.end());
  sig_max = 0.0;
  for (double sig : sigs) {
    if (sig > eps) {
      sig_max = sig;
      return false;
    } else if (sig > sig_max) {
      sig_max = sig;
    }
  }
  return true;
}

static void push_corner_pairs(std::vector<const tensor *> &all,
                              std::vector<tensor> *const A[4],
                              const std::vector<tensor> *old, int N) {
  for (int i = 0; i < N; ++i)
    for (int c = 0; c < 4; ++c) {
      all.push_back(&(*A[c])[i]);
      all.push_back(&old[c][i]);
    }
}

// ctm.cpp:815-969
bool Check_Convergence_CTM(const std::vector<tensor> &C1,
                           const std::vector<tensor> &C2,
                           const std::vector<tensor> &C3,
                           const std::vector<tensor> &C4,
                           const std::vector<tensor> &C1_old,
                           const std::vector<tensor> &C2_old,
                           const std::vector<tensor> &C3_old,
                           const std::vector<tensor> &C4_old,
                           const PEPS_Parameters &pp,
                           const SquareLattice &lat, double &sig_max,
                           double *sig_all) {
  const int N = lat.N_UNIT;
  std::vector<tensor> *const A[4] = {const_cast<std::vector<tensor> *>(&C1),
                                     const_cast<std::vector<tensor> *>(&C2),
                                     const_cast<std::vector<tensor> *>(&C3),
                                     const_cast<std::vector<tensor> *>(&C4)};
  const std::vector<tensor> old[4] = {C1_old, C2_old, C3_old, C4_old};
  std::vector<const tensor *> all;
  push_corner_pairs(all, A, old, N);
  const auto lam = spectra(all);
  return convergence_from_spectra(lam, 0, N, pp.CTM_Convergence_Epsilon, sig_max, sig_all);
}

namespace {

// One environment's CTM run inside run_ctm (the state of the reference's
// sweep loop, ctm.cpp:1196-1253).
struct CtmRun {
  std::vector<tensor> *A[8];
  const std::vector<tensor> *Tn;
  PEPS_Parameters pp;  // as given
  PEPS_Parameters ps;  // as run: the cut on the device when sweeps are captured
  const SquareLattice *lat;
  int64_t env = 0;
  bool graphs = false;
  bool convergence = false;
  int count = 0;
  double sig_max = 0.0;
  // corners of the previous sweep: handles, or copies when a replay may
  // write the current buffers in place
  std::vector<tensor> C_old[4];

  bool active() const { return !convergence && count < pp.Max_CTM_Iteration; }
  void keep_old() {
    for (int c = 0; c < 4; ++c) {
      C_old[c] = *A[c];
      if (graphs)
        for (auto &t : C_old[c]) t = tn::clone(t);
    }
  }
};

// The uncaptured sweep of one environment, in the reference's move order.
void eager_sweep(CtmRun &r) {
  auto &A = r.A;
  const auto &Tn = *r.Tn;
  const auto &lat = *r.lat;
  const auto &ps = r.ps;
  if (ps.CTM_Projector_corner) {
    corner_sweep(A, Tn, ps, lat);
    return;
  }
  auto &C1 = *A[0], &C2 = *A[1], &C3 = *A[2], &C4 = *A[3];
  auto &eTt = *A[4], &eTr = *A[5], &eTb = *A[6], &eTl = *A[7];
  const int LX = lat.LX_noskew, LY = lat.LY_noskew;
  for (int ix = 0; ix < LX; ++ix)
    Left_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, ix, ps, lat);
  for (int ix = 0; ix > -LX; --ix)
    Right_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, (ix + 1 + LX) % LX, ps, lat);
  for (int iy = 0; iy > -LY; --iy)
    Top_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, (iy + 1 + LY) % LY, ps, lat);
  for (int iy = 0; iy < LY; ++iy)
    Bottom_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, iy, ps, lat);
}

// The sweep loops of several environments of one context in lockstep: each
// round issues one sweep of every environment still running (uncaptured
// sweeps in turn, then the replays of the captured ones side by side, one
// per lane), then checks the convergence of all of them with one batched
// spectrum computation and one synchronisation.  Each environment's own
// sequence of sweeps and checks is the reference's (ctm.cpp:1196-1253).
void run_ctm(std::vector<CtmRun> &runs) {
  if (runs.empty()) return;
  tn::Context &ctx = *(*runs[0].Tn)[0].ctx();
  for (;;) {
    std::vector<CtmRun *> act;
    for (auto &r : runs)
      if (r.active()) act.push_back(&r);
    if (act.empty()) break;
    std::vector<const tn::Context::SweepGraph *> replays;
    for (CtmRun *r : act) {
      if (r->graphs && sweep_graph_ready(ctx, r->env, r->A, *r->lat, r->ps)) {
        auto it = ctx.sweep_graphs.find(r->env);
        if (it == ctx.sweep_graphs.end() ||
            it->second.signature != sweep_signature(ctx, r->env, r->A, *r->Tn, r->ps)) {
          ctx.drop_sweep_graph(r->env);
          capture_sweep(ctx, r->env, r->A, *r->Tn, r->ps, *r->lat);
          it = ctx.sweep_graphs.find(r->env);
        }
        if (it != ctx.sweep_graphs.end()) {
          replays.push_back(&it->second);
          continue;
        }
      }
      eager_sweep(*r);
    }
    if (!replays.empty()) {
      const int nl = std::min<int>((int)replays.size(), ctx.max_lanes());
      tn::ForkScope lanes(ctx, nl, nl > 1);
      for (size_t q = 0; q < replays.size(); ++q) {
        if (nl > 1) ctx.use_lane((int)q % nl);
        TN_HIP_CHECK(hipGraphLaunch(replays[q]->exec, ctx.stream()));
      }
      lanes.end();
      for (const auto *g : replays) {
        ctx.graph_replays++;
        ctx.pending_svd.insert(ctx.pending_svd.end(), g->pending.begin(), g->pending.end());
        ctx.gemm_flops += g->gemm_flops;
        ctx.gemm_calls += g->gemm_calls;
        ctx.svd_calls += g->svd_calls;
      }
    }
    // A run that cannot converge (epsilon <= 0: sig_max < epsilon never
    // holds) observes its checks only through the last one's sig_max (and
    // the per-sweep print at print level 3): when no active run needs this
    // round's check, it is skipped, and with it the round's synchronisation.
    // Runs with the cut on the device read their projector statistics at
    // that synchronisation and always take it.
    bool need_check = false;
    for (CtmRun *r : act)
      need_check = need_check || r->pp.CTM_Convergence_Epsilon > 0.0 ||
                   r->count + 1 >= r->pp.Max_CTM_Iteration || r->pp.print_level >= 3 ||
                   r->ps.projector_cut_on_device || r->graphs;
    if (!need_check) {
      for (CtmRun *r : act) {
        r->count += 1;
        r->keep_old();
      }
      continue;
    }
    std::vector<const tensor *> all;
    for (CtmRun *r : act) push_corner_pairs(all, r->A, r->C_old, r->lat->N_UNIT);
    const auto lam = spectra(all);
    // the spectra synchronised the stream: the sweeps' projector statistics
    // (and warm-start ranges) are in their pinned slots
    ctx.collect_svd_stats();
    size_t first = 0;
    for (CtmRun *r : act) {
      r->convergence = convergence_from_spectra(lam, first, r->lat->N_UNIT,
                                                r->pp.CTM_Convergence_Epsilon, r->sig_max, nullptr);
      first += 8 * (size_t)r->lat->N_UNIT;
      r->count += 1;
      r->keep_old();
      if (r->pp.print_level >= 3)
        std::printf("CTM: count, sig_max %d %.17g\n", r->count, r->sig_max);
    }
  }
  for (auto &r : runs)
    if (!r.convergence && r.pp.print_level >= 2)
      std::printf("Warning: CTM did not converge! count, sig_max = %d %.17g\n", r.count,
                  r.sig_max);
}

CtmRun make_run(CtmEnvironment &e, const PEPS_Parameters &pp, bool initialize) {
  tn::Context &ctx = *(*e.Tn)[0].ctx();
  CtmRun r;
  for (int k = 0; k < 8; ++k) r.A[k] = e.A[k];
  r.Tn = e.Tn;
  r.lat = e.lat;
  r.pp = pp;
  r.env = environment_id(*e.A[0]);
  // a fresh environment forgets the warm-start bases of its projectors (and
  // its captured sweep); a continued one keeps them
  if (initialize) {
    ctx.forget_svd_starts(r.env);
    Initialize_CTM(*e.A[0], *e.A[1], *e.A[2], *e.A[3], *e.A[4], *e.A[5], *e.A[6], *e.A[7],
                   *e.Tn, pp, *e.lat);
  }
  // captured sweeps need fixed shapes: the projector cut on the device
  r.ps = pp;
  r.graphs = pp.graph_sweeps && pp.CTM_Projector_corner && !ctx.profiling() &&
             !ctx.autotune() && !ctx.single_layer();
  if (r.graphs) r.ps.projector_cut_on_device = true;
  r.keep_old();
  return r;
}

}  // namespace

// ctm.cpp:971-1253
int Calc_CTM_Environment(std::vector<tensor> &C1, std::vector<tensor> &C2,
                         std::vector<tensor> &C3, std::vector<tensor> &C4,
                         std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                         std::vector<tensor> &eTb, std::vector<tensor> &eTl,
                         const std::vector<tensor> &Tn,
                         const PEPS_Parameters &pp, const SquareLattice &lat,
                         bool initialize, double *sig_max_out) {
  CtmEnvironment e{{&C1, &C2, &C3, &C4, &eTt, &eTr, &eTb, &eTl}, &Tn, &lat};
  std::vector<CtmRun> runs;
  runs.push_back(make_run(e, pp, initialize));
  run_ctm(runs);
  if (sig_max_out) *sig_max_out = runs[0].sig_max;
  return runs[0].count;
}

std::vector<int> Calc_CTM_Environment_lockstep(std::vector<CtmEnvironment> &envs,
                                               const std::vector<PEPS_Parameters> &pps,
                                               bool initialize, std::vector<double> *sig_max) {
  TN_REQUIRE(pps.size() == envs.size(), "lockstep CTM: one parameter set per environment");
  std::vector<CtmRun> runs;
  for (size_t i = 0; i < envs.size(); ++i) {
    TN_REQUIRE((*envs[i].Tn)[0].ctx() == (*envs[0].Tn)[0].ctx(),
               "lockstep CTM: the environments must share one context");
    runs.push_back(make_run(envs[i], pps[i], initialize));
  }
  run_ctm(runs);
  std::vector<int> counts;
  if (sig_max) sig_max->clear();
  for (auto &r : runs) {
    counts.push_back(r.count);
    if (sig_max) sig_max->push_back(r.sig_max);
  }
  return counts;
}

}  // namespace core
}  // namespace itps
}  // namespace tenes
