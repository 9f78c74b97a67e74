// Full update of one bond (full_update.hpp).
#include "full_update.hpp"

#include <algorithm>
#include <cmath>

#include "contract.hpp"
#include "small_dense.hpp"

namespace tenes {
namespace itps {

using tn::host::HT;
namespace H = tn::host;

namespace {

HT download(const tensor &t) {
  HT h(std::vector<int64_t>(t.shape.begin(), t.shape.end()));
  tn::to_host(t, h.v.data());
  return h;
}

tensor upload(tn::Context &ctx, const HT &h) {
  return tn::from_host(ctx, tn::Shape(h.shape.begin(), h.shape.end()), h.v.data());
}

// 1/s for s/s_max > cut, else 0 (core/full_update.cpp: the pseudo-inverse
// cut of LR1, LR2 and the ALS normal matrices)
std::vector<double> inv_vals(const std::vector<double> &s, double cut) {
  std::vector<double> r(s.size(), 0.0);
  const double smax = s.empty() ? 0.0 : s[0];
  for (size_t i = 0; i < s.size(); ++i)
    if (s[i] / smax > cut) r[i] = 1.0 / s[i];
  return r;
}

// pseudo-inverse in the form core/full_update.cpp builds it: (U s^-1) . VT
HT pinv_form(const HT &A, double cut) {
  HT U, VT;
  std::vector<double> s;
  H::svd(A, {0}, {1}, U, s, VT);
  H::scale_axis(U, 1, inv_vals(s, cut));
  return H::tensordot(U, VT, {1}, {0});
}

// -2 Re<Theta|R1 R2> + <R1 R2|R1 R2> in the metric Env (the ALS cost up to
// the constant <Theta|Theta>)
double als_delta(const HT &Env, const HT &Theta, const HT &R1, const HT &R2) {
  const double a = H::trace(
      H::tensordot(H::tensordot(Env, Theta, {0, 1}, {0, 1}), R2, {1, 3}, {0, 2}),
      R1, {0, 1, 2}, {0, 2, 1});
  const double b = H::trace(
      R1,
      H::tensordot(R2, H::tensordot(Env, H::tensordot(R1, R2, {1}, {1}), {2, 3}, {0, 2}),
                   {0, 2}, {1, 3}),
      {0, 1, 2}, {1, 0, 2});
  return -2.0 * a + b;
}

// One half-step of the ALS: the R that minimises the cost with the other R
// fixed, R = N^+ W (N, W as in core/full_update.cpp:320-366).
HT als_solve(const HT &N, const HT &W, double cut) {
  HT Uu, Vv;
  std::vector<double> s;
  H::svd(N, {1, 3}, {0, 2}, Uu, s, Vv);
  H::scale_axis(Uu, 2, inv_vals(s, cut));
  return H::tensordot(H::tensordot(Uu, Vv, {2}, {0}), W, {0, 1}, {0, 2});
}

}  // namespace

namespace core {

// core/full_update.cpp:85-393
void Full_update_bond_horizontal(const tensor &C1, const tensor &C2,
                                 const tensor &C3, const tensor &C4,
                                 const tensor &eT1, const tensor &eT2,
                                 const tensor &eT3, const tensor &eT4,
                                 const tensor &eT5, const tensor &eT6,
                                 const tensor &Tn1, const tensor &Tn2,
                                 const HT &op,
                                 const FU_Parameters &fp, tensor &Tn1_new,
                                 tensor &Tn2_new) {
  tn::Context &ctx = *Tn1.ctx();
  const int64_t D_connect = Tn1.shape[2];
  const int64_t d1 = Tn1.shape[4], d2 = Tn2.shape[4];
  TN_REQUIRE(op.rank() == 4 && op.shape[0] == d1 && op.shape[1] == d2 &&
                 op.size() == op.shape[0] * op.shape[1] * op.shape[2] * op.shape[3],
             "Full_update_bond: operator legs (d1, d2, d1o, d2o) do not match "
             "the physical dimensions of the site tensors");

  // QR of the site tensors: the bond and physical legs go to R
  HT Q1, R1, Q2, R2;
  H::qr(download(Tn1), {0, 1, 3}, {2, 4}, Q1, R1);
  H::qr(download(Tn2), {1, 2, 3}, {0, 4}, Q2, R2);
  const int64_t envR1 = R1.shape[0], envR2 = R2.shape[0];
  HT Theta = H::tensordot(H::tensordot(R1, R2, {1}, {1}), op, {1, 3}, {0, 1});

  // bond environment on the GPU (chi^3 D^4 contractions)
  HT Env = download(Create_Environment_two_sites(
      C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, upload(ctx, Q1),
      upload(ctx, H::transpose(Q2, {3, 0, 1, 2}))));
  Env = H::add(Env, H::transpose(Env, {2, 3, 0, 1}), 0.5, 0.5);

  // positive part of the environment: Env = Z Z^T
  std::vector<double> w, Zm;
  const int nenv = (int)(envR1 * envR2);
  H::eigh(Env.v, nenv, w, Zm);
  HT Z({envR1, envR2, (int64_t)nenv});
  Z.v = Zm;
  const double w_max = std::fabs(w[nenv - 1]);
  std::vector<double> ws(nenv, 0.0);
  for (int i = 0; i < nenv; ++i)
    if (w[i] / w_max > fp.Inverse_Env_cut) ws[i] = std::sqrt(std::max(w[i], 0.0));
  H::scale_axis(Z, 2, ws);

  HT LR1_inv, LR2_inv;
  if (fp.Full_Gauge_Fix) {
    HT q, LR1, LR2;
    H::qr(Z, {2, 1}, {0}, q, LR1);
    H::qr(Z, {2, 0}, {1}, q, LR2);
    Theta = H::tensordot(LR1, H::tensordot(LR2, Theta, {1}, {1}), {1}, {1});
    LR1_inv = pinv_form(LR1, fp.Inverse_Env_cut);
    LR2_inv = pinv_form(LR2, fp.Inverse_Env_cut);
    Z = H::tensordot(H::tensordot(Z, LR1_inv, {0}, {1}), LR2_inv, {0}, {1});
    Env = H::tensordot(Z, Z, {0}, {0});
  } else {
    Env = H::tensordot(Z, Z, {2}, {2});
  }

  // initial R1, R2 from the truncated SVD of Theta
  {
    HT U, VT;
    std::vector<double> s;
    H::svd(Theta, {0, 2}, {1, 3}, U, s, VT);
    const int64_t kc = std::min<int64_t>(D_connect, (int64_t)s.size());
    std::vector<double> lam(s.begin(), s.begin() + kc);
    double nrm = 0.0;
    for (double x : lam) nrm += x * x;
    nrm = std::sqrt(nrm);
    for (double &x : lam) x = std::sqrt(x / nrm);
    U = H::take(U, 2, kc);
    VT = H::take(VT, 0, kc);
    H::scale_axis(U, 2, lam);
    H::scale_axis(VT, 0, lam);
    R1 = H::transpose(U, {0, 2, 1});
    R2 = H::transpose(VT, {1, 0, 2});
  }
  const double C_phi = H::trace(H::tensordot(Env, Theta, {0, 1}, {0, 1}), Theta,
                                {0, 1, 2, 3}, {0, 1, 2, 3});
  double old = als_delta(Env, Theta, R1, R2);
  const HT EnvTheta = H::tensordot(Env, Theta, {0, 1}, {0, 1});
  bool conv = false;
  for (int count = 0; !conv && count < fp.Full_max_iteration; ++count) {
    HT W = H::tensordot(EnvTheta, R2, {1, 3}, {0, 2});
    HT N = H::tensordot(H::tensordot(Env, R2, {1}, {0}), R2, {2, 4}, {0, 2});
    R1 = als_solve(N, W, fp.Full_Inverse_precision);
    W = H::tensordot(EnvTheta, R1, {0, 2}, {0, 2});
    N = H::tensordot(H::tensordot(Env, R1, {0}, {0}), R1, {1, 4}, {0, 2});
    R2 = als_solve(N, W, fp.Full_Inverse_precision);
    const double delta = als_delta(Env, Theta, R1, R2);
    if (std::fabs(old - delta) / std::fabs(C_phi) < fp.Full_Convergence_Epsilon)
      conv = true;
    old = delta;
  }
  if (fp.Full_Gauge_Fix) {
    R1 = H::tensordot(LR1_inv, R1, {0}, {0});
    R2 = H::tensordot(LR2_inv, R2, {0}, {0});
  }

  // symmetric split of the new bond
  HT q1, r1, q2, r2, U, VT;
  H::qr(R1, {0, 2}, {1}, q1, r1);
  H::qr(R2, {0, 2}, {1}, q2, r2);
  std::vector<double> s;
  H::svd(H::tensordot(r1, r2, {1}, {1}), {0}, {1}, U, s, VT);
  const int64_t kc = std::min<int64_t>(D_connect, (int64_t)s.size());
  double nrm = 0.0;
  for (int64_t i = 0; i < kc; ++i) nrm += s[i] * s[i];
  nrm = std::sqrt(nrm);
  for (int64_t i = 0; i < kc; ++i) s[i] = std::sqrt(s[i] / nrm);
  H::scale_axis(U, 1, s);
  H::scale_axis(VT, 0, s);
  R1 = H::tensordot(q1, U, {2}, {0});
  R2 = H::tensordot(q2, VT, {2}, {1});
  Tn1_new = upload(ctx, H::transpose(H::tensordot(Q1, R1, {3}, {0}), {0, 1, 4, 2, 3}));
  Tn2_new = upload(ctx, H::transpose(H::tensordot(Q2, R2, {3}, {0}), {4, 0, 1, 2, 3}));
}

// core/full_update.cpp:395-441
void Full_update_bond(const tensor &C1, const tensor &C2, const tensor &C3,
                      const tensor &C4, const tensor &eT1, const tensor &eT2,
                      const tensor &eT3, const tensor &eT4, const tensor &eT5,
                      const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
                      const HT &op12, int connect1,
                      const FU_Parameters &fp, tensor &Tn1_new,
                      tensor &Tn2_new) {
  TN_REQUIRE(connect1 >= 0 && connect1 < 4, "Full_update_bond: bad leg");
  static const tn::Axes rot[4] = {{2, 3, 0, 1, 4}, {3, 0, 1, 2, 4}, {0, 1, 2, 3, 4},
                                  {1, 2, 3, 0, 4}};
  static const tn::Axes back[4] = {{2, 3, 0, 1, 4}, {1, 2, 3, 0, 4}, {0, 1, 2, 3, 4},
                                   {3, 0, 1, 2, 4}};
  const bool id = connect1 == 2;
  const tensor T1 = id ? Tn1 : tn::transpose(Tn1, rot[connect1]);
  const tensor T2 = id ? Tn2 : tn::transpose(Tn2, rot[connect1]);
  tensor n1, n2;
  Full_update_bond_horizontal(C1, C2, C3, C4, eT1, eT2, eT3, eT4, eT5, eT6, T1,
                              T2, op12, fp, n1, n2);
  Tn1_new = id ? n1 : tn::transpose(n1, back[connect1]);
  Tn2_new = id ? n2 : tn::transpose(n2, back[connect1]);
}

}  // namespace core

// src/iTPS/full_update.cpp:29-152
void full_update_step(std::vector<tensor> &C1, std::vector<tensor> &C2,
                      std::vector<tensor> &C3, std::vector<tensor> &C4,
                      std::vector<tensor> &eTt, std::vector<tensor> &eTr,
                      std::vector<tensor> &eTb, std::vector<tensor> &eTl,
                      std::vector<tensor> &Tn, int source, int leg,
                      const HT &op12,
                      const PEPS_Parameters &pp, const FU_Parameters &fp,
                      const SquareLattice &lat) {
  TN_REQUIRE(leg >= 0 && leg < 4, "full update: bad source leg");
  const int s = source, t = lat.neighbor(source, leg);
  // environment of the bond seen with the source on the left
  const tensor *a[10];
  if (leg == 0) {
    const tensor *x[10] = {&C4[s], &C2[t], &C1[t], &C3[s], &eTb[s],
                           &eTb[t], &eTl[t], &eTt[t], &eTt[s], &eTr[s]};
    std::copy(x, x + 10, a);
  } else if (leg == 1) {
    const tensor *x[10] = {&C4[s], &C1[t], &C2[t], &C3[s], &eTl[s],
                           &eTl[t], &eTt[t], &eTr[t], &eTr[s], &eTb[s]};
    std::copy(x, x + 10, a);
  } else if (leg == 2) {
    const tensor *x[10] = {&C1[s], &C2[t], &C3[t], &C4[s], &eTt[s],
                           &eTt[t], &eTr[t], &eTb[t], &eTb[s], &eTl[s]};
    std::copy(x, x + 10, a);
  } else {
    const tensor *x[10] = {&C2[s], &C3[t], &C4[t], &C1[s], &eTr[s],
                           &eTr[t], &eTb[t], &eTl[t], &eTl[s], &eTt[s]};
    std::copy(x, x + 10, a);
  }
  tensor n1, n2;
  core::Full_update_bond(*a[0], *a[1], *a[2], *a[3], *a[4], *a[5], *a[6], *a[7],
                         *a[8], *a[9], Tn[s], Tn[t], op12, leg, fp, n1, n2);
  Tn[s] = n1;
  Tn[t] = n2;
  if (!fp.Full_Use_FastFullUpdate) {
    core::Calc_CTM_Environment(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, pp, lat,
                               true, nullptr);
    return;
  }
  const int sx = s % lat.LX, tx = t % lat.LX, sy = s / lat.LX, ty = t / lat.LX;
  using namespace core;
  if (leg == 0) {
    Right_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, sx, pp, lat);
    Left_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, tx, pp, lat);
  } else if (leg == 1) {
    Bottom_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, sy, pp, lat);
    Top_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, ty, pp, lat);
  } else if (leg == 2) {
    Left_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, sx, pp, lat);
    Right_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, tx, pp, lat);
  } else {
    Top_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, sy, pp, lat);
    Bottom_move(C1, C2, C3, C4, eTt, eTr, eTb, eTl, Tn, ty, pp, lat);
  }
}

}  // namespace itps
}  // namespace tenes
