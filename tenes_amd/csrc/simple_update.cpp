// Simple update on the GPU.  Same steps as TeNeS core/simple_update.cpp:
// absorb the bond weights, QR both tensors, apply the gate to R1*R2, truncate
// the SVD to the bond dimension, split sqrt(s) to both sides and divide the
// weights back out.  QR is CholeskyQR (R with positive diagonal), the SVD the
// one-sided Jacobi of linalg.cpp; both differ from LAPACK only by a gauge on
// the new bond, which leaves every observable unchanged.
#include "simple_update.hpp"

#include <cmath>

namespace tenes {
namespace itps {
namespace core {

using tn::Axes;
using tn::Shape;

namespace {
// leg order that moves the connecting leg to position 3 (before phys), and
// the inverse placement of the new bond after the update
// (core/simple_update.cpp:97-120, 166-197)
const Axes kToEnd[4] = {{1, 2, 3, 0, 4}, {0, 2, 3, 1, 4}, {0, 1, 3, 2, 4},
                        {0, 1, 2, 3, 4}};
const Axes kBack1[4] = {{4, 0, 1, 2, 3}, {0, 4, 1, 2, 3}, {0, 1, 4, 2, 3},
                        {0, 1, 2, 4, 3}};
const Axes kBack2[4] = {{3, 0, 1, 2, 4}, {0, 3, 1, 2, 4}, {0, 1, 3, 2, 4},
                        {0, 1, 2, 3, 4}};
}  // namespace

void Simple_update_bond(const tensor &Tn1, const tensor &Tn2,
                        const std::vector<std::vector<double>> &lambda1,
                        const std::vector<std::vector<double>> &lambda2,
                        const tensor &op12, int connect1,
                        double inverse_lambda_cut, tensor &Tn1_new,
                        tensor &Tn2_new, std::vector<double> &lambda_c) {
  TN_REQUIRE(connect1 >= 0 && connect1 < 4, "simple update: bad leg");
  TN_REQUIRE(Tn1.rank() == 5 && Tn2.rank() == 5,
             "simple update: site tensors must have 5 legs");
  tn::Context &ctx = *Tn1.ctx();
  const int connect2 = (connect1 + 2) % 4;
  const int64_t dc = Tn1.shape[connect1];
  // one upload: weights of the three other legs of each site, then inverses
  std::vector<double> host;
  std::vector<int64_t> off1(4, -1), off2(4, -1), inv1(4, -1), inv2(4, -1);
  auto push = [&](const std::vector<double> &v, bool inverse) {
    const int64_t o = (int64_t)host.size();
    for (double x : v)
      host.push_back(inverse ? (x > inverse_lambda_cut ? 1.0 / x : 0.0) : x);
    return o;
  };
  for (int l = 0; l < 4; ++l) {
    if (l != connect1) {
      TN_REQUIRE((int64_t)lambda1[l].size() == Tn1.shape[l],
                 "simple update: lambda size mismatch");
      off1[l] = push(lambda1[l], false);
      inv1[l] = push(lambda1[l], true);
    }
    if (l != connect2) {
      TN_REQUIRE((int64_t)lambda2[l].size() == Tn2.shape[l],
                 "simple update: lambda size mismatch");
      off2[l] = push(lambda2[l], false);
      inv2[l] = push(lambda2[l], true);
    }
  }
  tensor lam = tn::from_host(ctx, Shape{(int64_t)host.size()}, host.data());
  tensor T1 = tn::clone(Tn1), T2 = tn::clone(Tn2);
  for (int l = 0; l < 4; ++l) {
    if (l != connect1) tn::multiply_vector_dev(T1, lam.data() + off1[l], l);
    if (l != connect2) tn::multiply_vector_dev(T2, lam.data() + off2[l], l);
  }
  T1 = tn::transpose(T1, kToEnd[connect1]);
  T2 = tn::transpose(T2, kToEnd[connect2]);
  // QR: rows = the three outer legs, cols = (connect, phys)
  auto qr = [&](const tensor &T, tensor &Q, tensor &R) {
    const int64_t m = T.shape[0] * T.shape[1] * T.shape[2];
    const int64_t n = T.shape[3] * T.shape[4];
    if (m < n) {
      // fewer rows than columns (a site whose physical leg a path-decomposed
      // gate has widened): the thin QR has k = m, and the new tensors
      // Q Uc do not depend on which orthonormal basis of R^m Q is (Q -> Q O
      // turns R into O^T R and the SVD factor Uc into O^T Uc), so Q = I,
      // R = T is exact
      std::vector<double> id((size_t)(m * m), 0.0);
      for (int64_t i = 0; i < m; ++i) id[(size_t)(i * m + i)] = 1.0;
      Q = tn::from_host(ctx, Shape{T.shape[0], T.shape[1], T.shape[2], m},
                        id.data());
      R = tn::reshape(T, Shape{m, T.shape[3], T.shape[4]});
      return;
    }
    tensor Rm;
    tensor Qm = tn::orthonormalize(ctx, tn::reshape(T, Shape{m, n}), &Rm);
    Q = tn::reshape(Qm, Shape{T.shape[0], T.shape[1], T.shape[2], n});
    R = tn::reshape(Rm, Shape{n, T.shape[3], T.shape[4]});
  };
  tensor Q1, R1, Q2, R2;
  qr(T1, Q1, R1);
  qr(T2, Q2, R2);
  // Theta (c1, c2, m1o, m2o), then matrix rows (c1, m1o), cols (c2, m2o)
  tensor Theta = tn::tensordot(tn::tensordot(R1, R2, Axes{1}, Axes{1}), op12,
                               Axes{1, 3}, Axes{0, 1});
  tensor Tm = tn::transpose(Theta, Axes{0, 2, 1, 3});
  const int64_t k1 = Theta.shape[0], k2 = Theta.shape[1];
  const int64_t m1 = Theta.shape[2], m2 = Theta.shape[3];
  const int kmax = (int)std::min(k1 * m1, k2 * m2);
  TN_REQUIRE(dc <= kmax, "simple update: bond dimension above the rank space");
  tn::SvdOptions so;
  tn::SvdResult sv =
      tn::svd_top(ctx, tn::reshape(Tm, Shape{k1 * m1, k2 * m2}), kmax, so);
  lambda_c.assign(sv.s.begin(), sv.s.begin() + dc);
  double norm = 0.0;
  for (double v : lambda_c) norm += v * v;
  norm = std::sqrt(norm);
  for (double &v : lambda_c) v = std::sqrt(v / norm);
  tensor lc = tn::from_host(ctx, Shape{dc}, lambda_c.data());
  tensor Uc = tn::slice(sv.U, 1, 0, dc);  // (k1*m1, dc)
  tensor Vc = tn::slice(sv.V, 1, 0, dc);  // (k2*m2, dc)
  tn::multiply_vector_dev(Uc, lc.data(), 1);
  tn::multiply_vector_dev(Vc, lc.data(), 1);
  Uc = tn::reshape(Uc, Shape{k1, m1, dc});
  Vc = tn::reshape(Vc, Shape{k2, m2, dc});  // = VT_c^T
  // divide the outer weights back out of Q1, Q2
  int q = 0;
  for (int l = 0; l < 4; ++l)
    if (l != connect1) tn::multiply_vector_dev(Q1, lam.data() + inv1[l], q++);
  q = 0;
  for (int l = 0; l < 4; ++l)
    if (l != connect2) tn::multiply_vector_dev(Q2, lam.data() + inv2[l], q++);
  // Tn1_new = (Q1 * Uc) (o0, o1, o2, m1, dc) -> put dc back on connect1
  Tn1_new = tn::tensordot_t(Q1, Uc, Axes{3}, Axes{0}, kBack1[connect1]);
  // Tn2_new = (Q2 * VTc) (o0, o1, o2, dc, m2) -> put dc back on connect2
  Tn2_new = tn::tensordot_t(Q2, Vc, Axes{3}, Axes{0},
                            [&] {
                              // result legs (o0,o1,o2,m2,dc): swap the last two
                              // then apply the reference's placement
                              Axes a = kBack2[connect2];
                              for (auto &x : a) x = (x == 3) ? 4 : (x == 4 ? 3 : x);
                              return a;
                            }());
}

void Simple_update_steps(std::vector<tensor> &Tn,
                         std::vector<std::vector<std::vector<double>>> &lambda,
                         const SquareLattice &lat,
                         const std::vector<EvolutionOp> &ops, int nsteps,
                         double inverse_lambda_cut) {
  for (int step = 0; step < nsteps; ++step) {
    for (const auto &up : ops) {
      const int source = up.source_site, leg = up.source_leg;
      const int target = lat.neighbor(source, leg);
      const int tleg = (leg + 2) % 4;
      tensor t1, t2;
      std::vector<double> lc;
      Simple_update_bond(Tn[source], Tn[target], lambda[source], lambda[target],
                         up.op, leg, inverse_lambda_cut, t1, t2, lc);
      lambda[source][leg] = lc;
      lambda[target][tleg] = lc;
      Tn[source] = t1;
      Tn[target] = t2;
    }
  }
}

}  // namespace core
}  // namespace itps
}  // namespace tenes
