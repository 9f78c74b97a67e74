// CTM-environment contractions (observables, full-update environment).
// Each follows the contraction tree of the TeNeS function it replaces; the
// final scalar is a device dot product read back once.
#include "contract.hpp"

#include <algorithm>
#include <sstream>

namespace tenes {
namespace itps {
namespace core {

using tn::Axes;
using tn::bra;
using tn::conj;
using tn::tensordot;
using tn::trace;

// contract_itps_ctm/ctm.cpp:89-113
double Contract_one_site_iTPS_CTM(const tensor &C1, const tensor &C2,
                                  const tensor &C3, const tensor &C4,
                                  const tensor &eT1, const tensor &eT2,
                                  const tensor &eT3, const tensor &eT4,
                                  const tensor &Tn1, const tensor &op1) {
  tensor upper =
      tensordot(tensordot(C2, tensordot(C3, eT2, Axes{0}, Axes{1}), Axes{1},
                          Axes{1}),
                tensordot(C1, eT1, Axes{1}, Axes{0}), Axes{0}, Axes{1});
  tensor lower = tensordot(
      tensordot(tensordot(C4, eT4, Axes{1}, Axes{0}), eT3, Axes{0}, Axes{1}),
      bra(Tn1), Axes{2, 5}, Axes{0, 3});
  tensor X = tensordot(tensordot(upper, lower, Axes{0, 2, 3, 5},
                                 Axes{2, 5, 0, 4}),
                       Tn1, Axes{0, 1, 2, 3}, Axes{2, 1, 0, 3});
  return trace(X, op1, Axes{0, 1}, Axes{1, 0});
}

// contract_itps_ctm/ctm.cpp:115-163
double Contract_two_sites_horizontal_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op1, const tensor &op2) {
  tensor right_half = tensordot(
      Tn2,
      tensordot(tensordot(tensordot(C3, eT4, Axes{1}, Axes{0}), eT3, Axes{0},
                          Axes{1}),
                tensordot(bra(Tn2), op2, Axes{4}, Axes{1}), Axes{2, 5},
                Axes{3, 2}),
      Axes{2, 3, 4}, Axes{3, 1, 6});
  tensor R = tensordot(
      tensordot(tensordot(C2, eT2, Axes{0}, Axes{1}), right_half,
                Axes{0, 2, 3}, Axes{3, 1, 5}),
      eT5, Axes{2}, Axes{0});
  tensor L = tensordot(tensordot(C1, eT1, Axes{1}, Axes{0}),
                       tensordot(C4, eT6, Axes{1}, Axes{0}), Axes{0}, Axes{1});
  tensor X = tensordot(
      Tn1,
      tensordot(L, tensordot(bra(Tn1), R, Axes{2, 3}, Axes{2, 5}),
                Axes{0, 2, 3, 5}, Axes{3, 1, 5, 0}),
      Axes{0, 1, 2, 3}, Axes{1, 0, 3, 4});
  return trace(X, op1, Axes{0, 1}, Axes{0, 1});
}

// contract_itps_ctm/ctm.cpp:165-211
double Contract_two_sites_vertical_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op1, const tensor &op2) {
  tensor lower_half = tensordot(
      Tn2,
      tensordot(tensordot(tensordot(C4, eT5, Axes{1}, Axes{0}), eT4, Axes{0},
                          Axes{1}),
                tensordot(bra(Tn2), op2, Axes{4}, Axes{1}), Axes{2, 5},
                Axes{0, 3}),
      Axes{0, 3, 4}, Axes{1, 3, 6});
  tensor B = tensordot(
      tensordot(tensordot(C3, eT3, Axes{0}, Axes{1}), lower_half,
                Axes{0, 2, 3}, Axes{3, 1, 5}),
      eT6, Axes{2}, Axes{0});
  tensor T = tensordot(tensordot(C2, eT2, Axes{1}, Axes{0}),
                       tensordot(C1, eT1, Axes{1}, Axes{0}), Axes{0}, Axes{1});
  tensor X = tensordot(
      Tn1,
      tensordot(T, tensordot(bra(Tn1), B, Axes{0, 3}, Axes{5, 2}),
                Axes{0, 2, 3, 5}, Axes{3, 1, 5, 0}),
      Axes{0, 1, 2, 3}, Axes{4, 1, 0, 3});
  return trace(X, op1, Axes{0, 1}, Axes{0, 1});
}

// contract_itps_ctm/ctm.cpp:213-266
double Contract_two_sites_horizontal_op12_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op12) {
  tensor left = tensordot(
      eT1,
      tensordot(Tn1,
                tensordot(bra(Tn1),
                          tensordot(eT5,
                                    tensordot(C1,
                                              tensordot(C4, eT6, Axes{1},
                                                        Axes{0}),
                                              Axes{0}, Axes{1}),
                                    Axes{1}, Axes{1}),
                          Axes{0, 3}, Axes{5, 2}),
                Axes{0, 3}, Axes{6, 4}),
      Axes{0, 2, 3}, Axes{7, 0, 3});
  tensor right = tensordot(
      eT2,
      tensordot(Tn2,
                tensordot(bra(Tn2),
                          tensordot(eT4,
                                    tensordot(C2,
                                              tensordot(C3, eT3, Axes{0},
                                                        Axes{1}),
                                              Axes{1}, Axes{1}),
                                    Axes{0}, Axes{1}),
                          Axes{2, 3}, Axes{5, 2}),
                Axes{2, 3}, Axes{6, 4}),
      Axes{1, 2, 3}, Axes{7, 1, 4});
  tensor X = tensordot(left, right, Axes{0, 1, 3, 5}, Axes{0, 1, 3, 5});
  return trace(op12, X, Axes{0, 1, 2, 3}, Axes{0, 2, 1, 3});
}

// contract_itps_ctm/ctm.cpp:268-319
double Contract_two_sites_vertical_op12_iTPS_CTM(
    const tensor &C1, const tensor &C2, const tensor &C3, const tensor &C4,
    const tensor &eT1, const tensor &eT2, const tensor &eT3, const tensor &eT4,
    const tensor &eT5, const tensor &eT6, const tensor &Tn1, const tensor &Tn2,
    const tensor &op12) {
  tensor top = tensordot(
      eT2,
      tensordot(Tn1,
                tensordot(bra(Tn1),
                          tensordot(eT6,
                                    tensordot(C1,
                                              tensordot(C2, eT1, Axes{0},
                                                        Axes{1}),
                                              Axes{1}, Axes{1}),
                                    Axes{1}, Axes{0}),
                          Axes{0, 1}, Axes{2, 5}),
                Axes{0, 1}, Axes{4, 6}),
      Axes{0, 2, 3}, Axes{7, 0, 3});
  tensor bottom = tensordot(
      eT3,
      tensordot(Tn2,
                tensordot(bra(Tn2),
                          tensordot(eT5,
                                    tensordot(C3,
                                              tensordot(C4, eT4, Axes{0},
                                                        Axes{1}),
                                              Axes{1}, Axes{1}),
                                    Axes{0}, Axes{1}),
                          Axes{0, 3}, Axes{2, 5}),
                Axes{0, 3}, Axes{4, 6}),
      Axes{1, 2, 3}, Axes{7, 1, 4});
  tensor X = tensordot(top, bottom, Axes{0, 1, 3, 5}, Axes{0, 1, 3, 5});
  return trace(op12, X, Axes{0, 1, 2, 3}, Axes{0, 2, 1, 3});
}

namespace {

// contract_itps_ctm/ctm_1x1.cpp:27-72
double contract_1x1(const std::vector<const tensor *> &C,
                    const std::vector<const tensor *> &eTt,
                    const std::vector<const tensor *> &eTr,
                    const std::vector<const tensor *> &eTb,
                    const std::vector<const tensor *> &eTl,
                    const std::vector<std::vector<const tensor *>> &Tn,
                    const std::vector<std::vector<const tensor *>> &op) {
  const tensor &T = *Tn[0][0];
  tensor top = tensordot(*eTt[0], tensordot(*C[0], *eTl[0], Axes{0}, Axes{1}),
                         Axes{0}, Axes{0});
  tensor rest = tensordot(
      tensordot(*C[1], *eTr[0], Axes{1}, Axes{0}),
      tensordot(*C[2], tensordot(*C[3], *eTb[0], Axes{0}, Axes{1}), Axes{1},
                Axes{1}),
      Axes{1}, Axes{0});
  tensor X = tensordot(
      T,
      tensordot(top, tensordot(bra(T), rest, Axes{2, 3}, Axes{2, 5}),
                Axes{0, 2, 3, 5}, Axes{3, 1, 5, 0}),
      Axes{0, 1, 2, 3}, Axes{1, 0, 3, 4});
  return trace(*op[0][0], X, Axes{0, 1}, Axes{0, 1});
}

// contract_itps_ctm/ctm_1x2.cpp:27-87
double contract_1x2(const std::vector<const tensor *> &C,
                    const std::vector<const tensor *> &eTt,
                    const std::vector<const tensor *> &eTr,
                    const std::vector<const tensor *> &eTb,
                    const std::vector<const tensor *> &eTl,
                    const std::vector<std::vector<const tensor *>> &Tn,
                    const std::vector<std::vector<const tensor *>> &op) {
  const tensor &T0 = *Tn[0][0];
  const tensor &T1 = *Tn[0][1];
  tensor right_env =
      tensordot(*eTb[1],
                tensordot(*C[1], tensordot(*C[2], *eTr[0], Axes{0}, Axes{1}),
                          Axes{1}, Axes{1}),
                Axes{0}, Axes{1});
  tensor site1 = tensordot(
      *eTt[1],
      tensordot(T1,
                tensordot(tensordot(bra(T1), *op[0][1], Axes{4}, Axes{1}),
                          right_env, Axes{2, 3}, Axes{5, 2}),
                Axes{2, 3, 4}, Axes{6, 4, 2}),
      Axes{1, 2, 3}, Axes{5, 1, 3});
  tensor lower = tensordot(tensordot(*C[3], *eTb[0], Axes{0}, Axes{1}), site1,
                           Axes{1}, Axes{3});
  tensor top = tensordot(*eTt[0], tensordot(*C[0], *eTl[0], Axes{0}, Axes{1}),
                         Axes{0}, Axes{0});
  tensor X = tensordot(
      T0,
      tensordot(top, tensordot(bra(T0), lower, Axes{2, 3}, Axes{5, 2}),
                Axes{0, 2, 3, 5}, Axes{5, 1, 3, 0}),
      Axes{0, 1, 2, 3}, Axes{1, 0, 4, 3});
  return trace(*op[0][0], X, Axes{0, 1}, Axes{0, 1});
}

// contract_itps_ctm/ctm_2x1.cpp:27-87
double contract_2x1(const std::vector<const tensor *> &C,
                    const std::vector<const tensor *> &eTt,
                    const std::vector<const tensor *> &eTr,
                    const std::vector<const tensor *> &eTb,
                    const std::vector<const tensor *> &eTl,
                    const std::vector<std::vector<const tensor *>> &Tn,
                    const std::vector<std::vector<const tensor *>> &op) {
  const tensor &T0 = *Tn[0][0];
  const tensor &T1 = *Tn[1][0];
  tensor bottom_env =
      tensordot(*eTl[1],
                tensordot(*C[2], tensordot(*C[3], *eTb[0], Axes{0}, Axes{1}),
                          Axes{1}, Axes{1}),
                Axes{0}, Axes{1});
  tensor site1 = tensordot(
      *eTr[1],
      tensordot(T1,
                tensordot(tensordot(bra(T1), *op[1][0], Axes{4}, Axes{1}),
                          bottom_env, Axes{0, 3}, Axes{2, 5}),
                Axes{0, 3, 4}, Axes{4, 6, 2}),
      Axes{1, 2, 3}, Axes{5, 1, 3});
  tensor right = tensordot(tensordot(*C[1], *eTr[0], Axes{1}, Axes{0}), site1,
                           Axes{1}, Axes{0});
  tensor top = tensordot(*eTt[0], tensordot(*C[0], *eTl[0], Axes{0}, Axes{1}),
                         Axes{0}, Axes{0});
  tensor X = tensordot(
      T0,
      tensordot(top, tensordot(bra(T0), right, Axes{2, 3}, Axes{2, 4}),
                Axes{0, 2, 3, 5}, Axes{3, 1, 6, 0}),
      Axes{0, 1, 2, 3}, Axes{1, 0, 3, 4});
  return trace(*op[0][0], X, Axes{0, 1}, Axes{0, 1});
}

// contract_itps_ctm/ctm_2x2.cpp:25-104
double contract_2x2(const std::vector<const tensor *> &C,
                    const std::vector<const tensor *> &eTt,
                    const std::vector<const tensor *> &eTr,
                    const std::vector<const tensor *> &eTb,
                    const std::vector<const tensor *> &eTl,
                    const std::vector<std::vector<const tensor *>> &Tn,
                    const std::vector<std::vector<const tensor *>> &op) {
  const tensor &T00 = *Tn[0][0], &T01 = *Tn[0][1], &T10 = *Tn[1][0],
               &T11 = *Tn[1][1];
  const tensor tr = tensordot(
      T01,
      tensordot(tensordot(bra(T01), *op[0][1], Axes{4}, Axes{1}),
                tensordot(*eTt[1], tensordot(*C[1], *eTr[0], Axes{1}, Axes{0}),
                          Axes{1}, Axes{0}),
                Axes{1, 2}, Axes{2, 5}),
      Axes{1, 2, 4}, Axes{4, 6, 2});
  const tensor bl = tensordot(
      T10,
      tensordot(tensordot(bra(T10), *op[1][0], Axes{4}, Axes{1}),
                tensordot(*eTb[0], tensordot(*C[3], *eTl[1], Axes{1}, Axes{0}),
                          Axes{1}, Axes{0}),
                Axes{0, 3}, Axes{5, 2}),
      Axes{0, 3, 4}, Axes{6, 4, 2});
  const tensor br = tensordot(
      T11,
      tensordot(tensordot(bra(T11), *op[1][1], Axes{4}, Axes{1}),
                tensordot(*eTr[1], tensordot(*C[2], *eTb[1], Axes{1}, Axes{0}),
                          Axes{1}, Axes{0}),
                Axes{2, 3}, Axes{2, 5}),
      Axes{2, 3, 4}, Axes{4, 6, 2});
  const tensor rest = tensordot(tr, tensordot(bl, br, Axes{1, 3, 4}, Axes{0, 2, 5}),
                                Axes{1, 3, 5}, Axes{3, 4, 5});
  const tensor top = tensordot(*eTt[0], tensordot(*C[0], *eTl[0], Axes{0}, Axes{1}),
                               Axes{0}, Axes{0});
  const tensor X = tensordot(
      T00,
      tensordot(top, tensordot(bra(T00), rest, Axes{2, 3}, Axes{1, 4}),
                Axes{0, 2, 3, 5}, Axes{4, 1, 6, 0}),
      Axes{0, 1, 2, 3}, Axes{1, 0, 3, 4});
  return trace(*op[0][0], X, Axes{0, 1}, Axes{0, 1});
}


// A tensor with one integer label per leg.
struct Labeled {
  tensor t;
  std::vector<int> lab;
};

// Contracts every label the operands share; the result carries the free
// legs of A, then those of B (the tensordot order).
Labeled dot(const Labeled &A, const Labeled &B) {
  Axes ax, bx;
  std::vector<int> out;
  for (size_t i = 0; i < A.lab.size(); ++i) {
    auto it = std::find(B.lab.begin(), B.lab.end(), A.lab[i]);
    if (it == B.lab.end()) {
      out.push_back(A.lab[i]);
    } else {
      ax.push_back((int)i);
      bx.push_back((int)(it - B.lab.begin()));
    }
  }
  for (int l : B.lab)
    if (std::find(A.lab.begin(), A.lab.end(), l) == A.lab.end())
      out.push_back(l);
  TN_REQUIRE(!ax.empty(), "cluster sweep: operands share no leg");
  return {tensordot(A.t, B.t, ax, bx), out};
}

// Full contraction of two tensors carrying the same labels.
double close(const Labeled &A, const Labeled &B) {
  TN_REQUIRE(A.lab.size() == B.lab.size(), "cluster sweep: open legs left");
  Axes ax, bx;
  for (size_t i = 0; i < A.lab.size(); ++i) {
    auto it = std::find(B.lab.begin(), B.lab.end(), A.lab[i]);
    TN_REQUIRE(it != B.lab.end(), "cluster sweep: open legs left");
    ax.push_back((int)i);
    bx.push_back((int)(it - B.lab.begin()));
  }
  return trace(A.t, B.t, ax, bx);
}

}  // namespace

// Leg labels of the cluster network (C/eT leg orders of ctm.cpp:984-1192,
// site legs (l, u, r, d, p)); the operator's first index meets the ket's
// physical leg, its second the bra's, as in the ctm_NxM.cpp trees.
double contract_cluster_sweep(const std::vector<const tensor *> &C,
                              const std::vector<const tensor *> &eTt,
                              const std::vector<const tensor *> &eTr,
                              const std::vector<const tensor *> &eTb,
                              const std::vector<const tensor *> &eTl,
                              const std::vector<std::vector<const tensor *>> &Tn,
                              const std::vector<std::vector<const tensor *>> &op) {
  const int nr = (int)Tn.size();
  const int nc = nr ? (int)Tn[0].size() : 0;
  TN_REQUIRE(nr > 0 && nc > 0 && C.size() == 4 && (int)eTt.size() == nc &&
                 (int)eTb.size() == nc && (int)eTl.size() == nr &&
                 (int)eTr.size() == nr && (int)op.size() == nr,
             "cluster sweep: inconsistent cluster");
  // ring bonds: top chain top(0..nc), right chain rgt(0..nr), bottom chain
  // bot(0..nc) (bot(0) at C4), left chain lft(0..nr) (lft(0) at C1); ket /
  // bra bonds: horizontal h(r, 0..nc), vertical v(0..nr, c); physical p(r, c)
  const int nh = nr * (nc + 1), nv = (nr + 1) * nc;
  auto top = [&](int c) { return c; };
  auto rgt = [&](int r) { return nc + 1 + r; };
  auto bot = [&](int c) { return nc + nr + 2 + c; };
  auto lft = [&](int r) { return 2 * nc + nr + 3 + r; };
  const int base = 2 * (nc + nr) + 4;
  auto hk = [&](int r, int c) { return base + r * (nc + 1) + c; };
  auto hb = [&](int r, int c) { return base + nh + r * (nc + 1) + c; };
  auto vk = [&](int r, int c) { return base + 2 * nh + r * nc + c; };
  auto vb = [&](int r, int c) { return base + 2 * nh + nv + r * nc + c; };
  auto ph = [&](int r, int c) { return base + 2 * nh + 2 * nv + r * nc + c; };

  const Labeled c1{*C[0], {lft(0), top(0)}};
  const Labeled c2{*C[1], {top(nc), rgt(0)}};
  const Labeled c3{*C[2], {rgt(nr), bot(nc)}};
  const Labeled c4{*C[3], {bot(0), lft(nr)}};
  auto et = [&](int c) {
    return Labeled{*eTt[c], {top(c), top(c + 1), vk(0, c), vb(0, c)}};
  };
  auto er = [&](int r) {
    return Labeled{*eTr[r], {rgt(r), rgt(r + 1), hk(r, nc), hb(r, nc)}};
  };
  auto eb = [&](int c) {
    return Labeled{*eTb[c], {bot(c + 1), bot(c), vk(nr, c), vb(nr, c)}};
  };
  auto el = [&](int r) {
    return Labeled{*eTl[r], {lft(r + 1), lft(r), hk(r, 0), hb(r, 0)}};
  };
  auto ket = [&](int r, int c) {
    return Labeled{*Tn[r][c],
                   {hk(r, c), vk(r, c), hk(r, c + 1), vk(r + 1, c), ph(r, c)}};
  };
  auto bra_of = [&](int r, int c) {
    TN_REQUIRE((int)op[r].size() == nc && op[r][c] && Tn[r][c],
               "cluster sweep: missing site tensor or operator");
    return Labeled{tensordot(bra(*Tn[r][c]), *op[r][c], Axes{4}, Axes{1}),
                   {hb(r, c), vb(r, c), hb(r, c + 1), vb(r + 1, c), ph(r, c)}};
  };

  if (nr <= nc) {  // columns left to right; the boundary spans nr rows
    Labeled B = c1;
    for (int r = 0; r < nr; ++r) B = dot(B, el(r));
    B = dot(B, c4);
    for (int c = 0; c < nc; ++c) {
      B = dot(B, et(c));
      for (int r = 0; r < nr; ++r) {
        B = dot(B, ket(r, c));
        B = dot(B, bra_of(r, c));
      }
      B = dot(B, eb(c));
    }
    B = dot(B, c2);
    for (int r = 0; r < nr; ++r) B = dot(B, er(r));
    return close(B, c3);
  }
  // rows top to bottom; the boundary spans nc columns
  Labeled B = c1;
  for (int c = 0; c < nc; ++c) B = dot(B, et(c));
  B = dot(B, c2);
  for (int r = 0; r < nr; ++r) {
    B = dot(B, el(r));
    for (int c = 0; c < nc; ++c) {
      B = dot(B, ket(r, c));
      B = dot(B, bra_of(r, c));
    }
    B = dot(B, er(r));
  }
  B = dot(B, c4);
  for (int c = 0; c < nc; ++c) B = dot(B, eb(c));
  return close(B, c3);
}

double Contract_iTPS_CTM(const std::vector<const tensor *> &C,
                         const std::vector<const tensor *> &eTt,
                         const std::vector<const tensor *> &eTr,
                         const std::vector<const tensor *> &eTb,
                         const std::vector<const tensor *> &eTl,
                         const std::vector<std::vector<const tensor *>> &Tn,
                         const std::vector<std::vector<const tensor *>> &op) {
  const size_t nrow = Tn.size();
  const size_t ncol = nrow ? Tn[0].size() : 0;
  if (nrow == 1 && ncol == 1) return contract_1x1(C, eTt, eTr, eTb, eTl, Tn, op);
  if (nrow == 1 && ncol == 2) return contract_1x2(C, eTt, eTr, eTb, eTl, Tn, op);
  if (nrow == 2 && ncol == 1) return contract_2x1(C, eTt, eTr, eTb, eTl, Tn, op);
  if (nrow == 2 && ncol == 2) return contract_2x2(C, eTt, eTr, eTb, eTl, Tn, op);
  if (nrow >= 1 && nrow <= 4 && ncol >= 1 && ncol <= 4)
    return contract_cluster_sweep(C, eTt, eTr, eTb, eTl, Tn, op);
  std::ostringstream ss;
  ss << "Contract_" << nrow << "_" << ncol << "_iTPS_CTM is not implemented";
  throw tn::Error(ss.str(), -5);
}

// core/full_update.cpp:51-83
tensor Create_Environment_two_sites(const tensor &C1, const tensor &C2,
                                    const tensor &C3, const tensor &C4,
                                    const tensor &eT1, const tensor &eT2,
                                    const tensor &eT3, const tensor &eT4,
                                    const tensor &eT5, const tensor &eT6,
                                    const tensor &Q1, const tensor &Q2) {
  tensor right_q = tensordot(
      tensordot(tensordot(tensordot(C3, eT4, Axes{1}, Axes{0}), eT3, Axes{0},
                          Axes{1}),
                conj(Q2), Axes{2, 5}, Axes{3, 2}),
      Q2, Axes{1, 3}, Axes{3, 2});
  tensor right = tensordot(tensordot(C2, eT2, Axes{0}, Axes{1}), right_q,
                           Axes{0, 2, 3}, Axes{1, 5, 3});
  tensor left_q = tensordot(
      tensordot(tensordot(tensordot(C4, eT6, Axes{1}, Axes{0}), eT5, Axes{0},
                          Axes{1}),
                conj(Q1), Axes{2, 5}, Axes{0, 2}),
      Q1, Axes{1, 3}, Axes{0, 2});
  tensor left = tensordot(tensordot(C1, eT1, Axes{1}, Axes{0}), left_q,
                          Axes{0, 2, 3}, Axes{0, 4, 2});
  return tn::tensordot_t(right, left, Axes{0, 1}, Axes{0, 1},
                         Axes{3, 1, 2, 0});
}

// contract_itps_ctm/correlation.cpp:34-58
void StartCorrelation_iTPS_CTM(tensor &A, const tensor &C1, const tensor &C4,
                               const tensor &eT1, const tensor &eT3,
                               const tensor &eT4, const tensor &Tn1,
                               const tensor &op) {
  const tensor inner = tensordot(
      tensordot(bra(Tn1), op, Axes{4}, Axes{1}),
      tensordot(eT3, tensordot(C1, tensordot(C4, eT4, Axes{1}, Axes{0}),
                               Axes{0}, Axes{1}),
                Axes{1}, Axes{1}),
      Axes{0, 3}, Axes{5, 2});
  A = tn::tensordot_t(eT1, tensordot(Tn1, inner, Axes{0, 3, 4}, Axes{6, 4, 2}),
                      Axes{0, 2, 3}, Axes{5, 0, 2}, Axes{0, 3, 1, 2});
}

// contract_itps_ctm/correlation.cpp:60-78
void Transfer_iTPS_CTM(tensor &A, const tensor &eT1, const tensor &eT3,
                       const tensor &Tn1) {
  const tensor inner = tensordot(bra(Tn1), tensordot(A, eT3, Axes{1}, Axes{1}),
                                 Axes{0, 3}, Axes{2, 5});
  A = tn::tensordot_t(eT1, tensordot(Tn1, inner, Axes{0, 3, 4}, Axes{4, 6, 2}),
                      Axes{0, 2, 3}, Axes{4, 0, 2}, Axes{0, 3, 1, 2});
}

// contract_itps_ctm/correlation.cpp:80-105
double FinishCorrelation_iTPS_CTM(const tensor &A, const tensor &C2,
                                  const tensor &C3, const tensor &eT1,
                                  const tensor &eT2, const tensor &eT3,
                                  const tensor &Tn1, const tensor &op) {
  const tensor right = tensordot(
      bra(Tn1),
      tensordot(eT3, tensordot(C2, tensordot(C3, eT2, Axes{0}, Axes{1}),
                               Axes{1}, Axes{1}),
                Axes{0}, Axes{1}),
      Axes{2, 3}, Axes{5, 2});
  const tensor X = tensordot(
      Tn1,
      tensordot(tensordot(A, eT1, Axes{0}, Axes{0}), right, Axes{0, 2, 3, 5},
                Axes{3, 0, 5, 1}),
      Axes{0, 1, 2, 3}, Axes{0, 1, 4, 3});
  return trace(op, X, Axes{0, 1}, Axes{0, 1});
}

}  // namespace core
}  // namespace itps
}  // namespace tenes
