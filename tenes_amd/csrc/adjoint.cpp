// Bond reduced density matrices and their adjoint (see adjoint.hpp).
#include "adjoint.hpp"

#include <algorithm>

#include "kernels.hpp"

namespace tenes {
namespace itps {
namespace core {

using tn::Axes;
using tn::tensordot;

namespace {

Axes free_legs(int rank, const Axes &ax) {
  Axes out;
  for (int i = 0; i < rank; ++i)
    if (std::find(ax.begin(), ax.end(), i) == ax.end()) out.push_back(i);
  return out;
}

int rank_in_sorted(const Axes &ax, int v) {
  int r = 0;
  for (int a : ax) r += a < v;
  return r;
}

// For C = tensordot(A, B, axA, axB) and dC = d s / dC: d s / dA is
// tensordot(dC, B) over B's free legs, whose result carries A's free legs
// followed by B's contracted legs in ascending order; `perm` puts them back
// into A's leg order.  G = dA + beta * G.
void grad_lhs_into(const tensor &dC, const tensor &A, const tensor &B,
                   const Axes &axA, const Axes &axB, tensor &G, double beta) {
  const int na = A.rank(), nb = B.rank();
  const Axes fa = free_legs(na, axA), fb = free_legs(nb, axB);
  const int nfa = (int)fa.size();
  TN_REQUIRE(!fb.empty(), "adjoint: right operand without free legs");
  Axes dax(fb.size());
  for (size_t i = 0; i < fb.size(); ++i) dax[i] = nfa + (int)i;
  Axes perm(na);
  for (int i = 0; i < nfa; ++i) perm[fa[i]] = i;
  for (size_t k = 0; k < axA.size(); ++k)
    perm[axA[k]] = nfa + rank_in_sorted(axB, axB[k]);
  tn::tensordot_into(dC, B, dax, fb, perm, G, 1.0, beta);
}

// d s / dB = tensordot(A, dC) over A's free legs: A's contracted legs in
// ascending order, then B's free legs.
tensor grad_rhs(const tensor &dC, const tensor &A, const tensor &B,
                const Axes &axA, const Axes &axB) {
  const int na = A.rank(), nb = B.rank();
  const Axes fa = free_legs(na, axA), fb = free_legs(nb, axB);
  const int nfa = (int)fa.size(), nc = (int)axA.size();
  TN_REQUIRE(nfa > 0, "adjoint: left operand without free legs");
  Axes dax(nfa);
  for (int i = 0; i < nfa; ++i) dax[i] = i;
  Axes perm(nb);
  for (size_t i = 0; i < fb.size(); ++i) perm[fb[i]] = nc + (int)i;
  for (int k = 0; k < nc; ++k) perm[axB[k]] = rank_in_sorted(axA, axA[k]);
  return tn::tensordot_t(A, dC, fa, dax, perm);
}

void finish_half(BondHalf &h) {
  h.a4 = tensordot(tn::conj(*h.T), h.a3, h.ax4A, h.ax4B);
  h.a5 = tensordot(*h.T, h.a4, h.ax5A, h.ax5B);
  h.full = tensordot(h.e, h.a5, h.ax6A, h.ax6B);
}

void half_adjoint(const BondHalf &h, const tensor &dfull, tensor &G) {
  const tensor da5 = grad_rhs(dfull, h.e, h.a5, h.ax6A, h.ax6B);
  grad_lhs_into(da5, *h.T, h.a4, h.ax5A, h.ax5B, G, 1.0);
  const tensor da4 = grad_rhs(da5, *h.T, h.a4, h.ax5A, h.ax5B);
  grad_lhs_into(da4, tn::conj(*h.T), h.a3, h.ax4A, h.ax4B, G, 1.0);
}

const Axes kJoin{0, 1, 3, 5};

}  // namespace

// contract_itps_ctm/ctm.cpp:213-266 up to the operator
BondTape bond_rdm_horizontal(const tensor &C1, const tensor &C2,
                             const tensor &C3, const tensor &C4,
                             const tensor &eT1, const tensor &eT2,
                             const tensor &eT3, const tensor &eT4,
                             const tensor &eT5, const tensor &eT6,
                             const tensor &Tn1, const tensor &Tn2) {
  BondTape t;
  BondHalf &l = t.h1;
  l.T = &Tn1;
  l.e = eT1;
  l.a3 = tensordot(eT5, tensordot(C1, tensordot(C4, eT6, Axes{1}, Axes{0}),
                                  Axes{0}, Axes{1}),
                   Axes{1}, Axes{1});
  l.ax4A = {0, 3};
  l.ax4B = {5, 2};
  l.ax5A = {0, 3};
  l.ax5B = {6, 4};
  l.ax6A = {0, 2, 3};
  l.ax6B = {7, 0, 3};
  finish_half(l);
  BondHalf &r = t.h2;
  r.T = &Tn2;
  r.e = eT2;
  r.a3 = tensordot(eT4, tensordot(C2, tensordot(C3, eT3, Axes{0}, Axes{1}),
                                  Axes{1}, Axes{1}),
                   Axes{0}, Axes{1});
  r.ax4A = {2, 3};
  r.ax4B = {5, 2};
  r.ax5A = {2, 3};
  r.ax5B = {6, 4};
  r.ax6A = {1, 2, 3};
  r.ax6B = {7, 1, 4};
  finish_half(r);
  t.rho = tensordot(l.full, r.full, kJoin, kJoin);
  return t;
}

// contract_itps_ctm/ctm.cpp:268-319 up to the operator
BondTape bond_rdm_vertical(const tensor &C1, const tensor &C2,
                           const tensor &C3, const tensor &C4,
                           const tensor &eT1, const tensor &eT2,
                           const tensor &eT3, const tensor &eT4,
                           const tensor &eT5, const tensor &eT6,
                           const tensor &Tn1, const tensor &Tn2) {
  BondTape t;
  BondHalf &u = t.h1;
  u.T = &Tn1;
  u.e = eT2;
  u.a3 = tensordot(eT6, tensordot(C1, tensordot(C2, eT1, Axes{0}, Axes{1}),
                                  Axes{1}, Axes{1}),
                   Axes{1}, Axes{0});
  u.ax4A = {0, 1};
  u.ax4B = {2, 5};
  u.ax5A = {0, 1};
  u.ax5B = {4, 6};
  u.ax6A = {0, 2, 3};
  u.ax6B = {7, 0, 3};
  finish_half(u);
  BondHalf &b = t.h2;
  b.T = &Tn2;
  b.e = eT3;
  b.a3 = tensordot(eT5, tensordot(C3, tensordot(C4, eT4, Axes{0}, Axes{1}),
                                  Axes{1}, Axes{1}),
                   Axes{0}, Axes{1});
  b.ax4A = {0, 3};
  b.ax4B = {2, 5};
  b.ax5A = {0, 3};
  b.ax5B = {4, 6};
  b.ax6A = {1, 2, 3};
  b.ax6B = {7, 1, 4};
  finish_half(b);
  t.rho = tensordot(u.full, b.full, kJoin, kJoin);
  return t;
}

void bond_rdm_adjoint(const BondTape &tape, const tensor &dRho, tensor &G1,
                      tensor &G2) {
  // rho = tensordot(h1.full, h2.full, kJoin, kJoin)
  tensor d1 = tn::empty(*dRho.ctx(), tape.h1.full.shape);
  grad_lhs_into(dRho, tape.h1.full, tape.h2.full, kJoin, kJoin, d1, 0.0);
  const tensor d2 = grad_rhs(dRho, tape.h1.full, tape.h2.full, kJoin, kJoin);
  half_adjoint(tape.h1, d1, G1);
  half_adjoint(tape.h2, d2, G2);
}

double bond_energy(const std::vector<tensor> *env,
                   const std::vector<BondTerm> &terms, int n_sites,
                   std::vector<tensor> *G) {
  TN_REQUIRE(n_sites > 0, "bond_energy: no sites");
  const std::vector<tensor> &Tn = env[8];
  if (G) {
    TN_REQUIRE((int)G->size() == n_sites, "bond_energy: one gradient per site");
    for (int i = 0; i < n_sites; ++i) {
      tensor &g = (*G)[i];
      TN_REQUIRE(g.shape == Tn[i].shape,
                 "bond_energy: gradient shape differs from the site tensor");
      tn::kern::fill(g.data(), g.size(), 0.0, g.ctx()->stream());
    }
  }
  double E = 0.0;
  for (const BondTerm &b : terms) {
    const int s1 = b.s1, s2 = b.s2;
    TN_REQUIRE(s1 >= 0 && s1 < n_sites && s2 >= 0 && s2 < n_sites,
               "bond_energy: bad site index");
    const int64_t d1 = Tn[s1].shape[4], d2 = Tn[s2].shape[4];
    TN_REQUIRE((int64_t)b.op.size() == d1 * d2 * d1 * d2,
               "bond_energy: operator size differs from d1*d2*d1*d2");
    const BondTape tape =
        b.vertical
            ? bond_rdm_vertical(env[0][s1], env[1][s1], env[2][s2], env[3][s2],
                                env[4][s1], env[5][s1], env[5][s2], env[6][s2],
                                env[7][s2], env[7][s1], Tn[s1], Tn[s2])
            : bond_rdm_horizontal(env[0][s1], env[1][s2], env[2][s2],
                                  env[3][s1], env[4][s1], env[4][s2],
                                  env[5][s2], env[6][s2], env[6][s1],
                                  env[7][s1], Tn[s1], Tn[s2]);
    std::vector<double> rho(d1 * d1 * d2 * d2);
    tn::to_host(tape.rho, rho.data());
    auto R = [&](int64_t p, int64_t q, int64_t t, int64_t u) {
      return rho[((p * d1 + q) * d2 + t) * d2 + u];
    };
    auto O = [&](int64_t p, int64_t t, int64_t q, int64_t u) {
      return b.op[((p * d2 + t) * d1 + q) * d2 + u];
    };
    double num = 0.0, den = 0.0;
    for (int64_t p = 0; p < d1; ++p)
      for (int64_t q = 0; q < d1; ++q)
        for (int64_t t = 0; t < d2; ++t)
          for (int64_t u = 0; u < d2; ++u) {
            num += O(p, t, q, u) * R(p, q, t, u);
            if (p == q && t == u) den += R(p, q, t, u);
          }
    TN_REQUIRE(den != 0.0, "bond_energy: zero norm");
    const double Eb = num / den;
    E += Eb;
    if (G) {
      std::vector<double> drho(rho.size());
      const double w = 1.0 / (den * n_sites);
      for (int64_t p = 0; p < d1; ++p)
        for (int64_t q = 0; q < d1; ++q)
          for (int64_t t = 0; t < d2; ++t)
            for (int64_t u = 0; u < d2; ++u)
              drho[((p * d1 + q) * d2 + t) * d2 + u] =
                  (O(p, t, q, u) - (p == q && t == u ? Eb : 0.0)) * w;
      const tensor dRho =
          tn::from_host(*tape.rho.ctx(), tape.rho.shape, drho.data());
      bond_rdm_adjoint(tape, dRho, (*G)[s1], (*G)[s2]);
    }
  }
  if (G) Tn[0].ctx()->sync();
  return E / n_sites;
}

}  // namespace core
}  // namespace itps
}  // namespace tenes
