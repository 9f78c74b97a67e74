// Small dense FP64 tensors and factorizations on the host (small_dense.hpp).
#include "small_dense.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "common.hpp"

namespace tn {
namespace host {

namespace {

int64_t prod(const std::vector<int64_t> &s, size_t b = 0, size_t e = SIZE_MAX) {
  int64_t p = 1;
  for (size_t i = b; i < std::min(e, s.size()); ++i) p *= s[i];
  return p;
}

std::vector<int> free_legs(int rank, const std::vector<int> &ax) {
  std::vector<int> out;
  for (int i = 0; i < rank; ++i)
    if (std::find(ax.begin(), ax.end(), i) == ax.end()) out.push_back(i);
  return out;
}

}  // namespace

HT::HT(std::vector<int64_t> s) : shape(std::move(s)), v(prod(shape), 0.0) {}

HT transpose(const HT &t, const std::vector<int> &perm) {
  const int r = t.rank();
  TN_REQUIRE((int)perm.size() == r, "host transpose: wrong permutation size");
  std::vector<int64_t> ns(r), ostride(r), step(r);
  int64_t s = 1;
  for (int i = r - 1; i >= 0; --i) {
    ostride[i] = s;
    s *= t.shape[i];
  }
  for (int k = 0; k < r; ++k) {
    ns[k] = t.shape[perm[k]];
    step[k] = ostride[perm[k]];
  }
  HT out(ns);
  std::vector<int64_t> idx(r, 0);
  int64_t off = 0;
  for (int64_t n = 0; n < out.size(); ++n) {
    out.v[n] = t.v[off];
    for (int k = r - 1; k >= 0; --k) {  // odometer over the new shape
      if (++idx[k] < ns[k]) {
        off += step[k];
        break;
      }
      off -= step[k] * (ns[k] - 1);
      idx[k] = 0;
    }
  }
  return out;
}

HT reshape(HT t, std::vector<int64_t> shape) {
  TN_REQUIRE(prod(shape) == t.size(), "host reshape: size mismatch");
  t.shape = std::move(shape);
  return t;
}

HT tensordot(const HT &A, const HT &B, const std::vector<int> &axA,
             const std::vector<int> &axB) {
  TN_REQUIRE(axA.size() == axB.size(), "host tensordot: axis count mismatch");
  for (size_t k = 0; k < axA.size(); ++k)
    TN_REQUIRE(A.shape[axA[k]] == B.shape[axB[k]],
               "host tensordot: contracted dimensions differ");
  const std::vector<int> fa = free_legs(A.rank(), axA), fb = free_legs(B.rank(), axB);
  std::vector<int> pa = fa, pb = axB;
  pa.insert(pa.end(), axA.begin(), axA.end());
  pb.insert(pb.end(), fb.begin(), fb.end());
  const HT At = transpose(A, pa), Bt = transpose(B, pb);
  int64_t M = 1, N = 1, K = 1;
  std::vector<int64_t> shape;
  for (int i : fa) {
    M *= A.shape[i];
    shape.push_back(A.shape[i]);
  }
  for (int i : fb) {
    N *= B.shape[i];
    shape.push_back(B.shape[i]);
  }
  for (int i : axA) K *= A.shape[i];
  HT C(shape);
  for (int64_t i = 0; i < M; ++i)
    for (int64_t k = 0; k < K; ++k) {
      const double a = At.v[i * K + k];
      if (a == 0.0) continue;
      const double *b = &Bt.v[k * N];
      double *c = &C.v[i * N];
      for (int64_t j = 0; j < N; ++j) c[j] += a * b[j];
    }
  return C;
}

double trace(const HT &A, const HT &B, const std::vector<int> &axA,
             const std::vector<int> &axB) {
  TN_REQUIRE((int)axA.size() == A.rank() && (int)axB.size() == B.rank() &&
                 A.size() == B.size(),
             "host trace: every leg must be contracted");
  std::vector<int> perm(B.rank());
  for (size_t k = 0; k < axA.size(); ++k) perm[axA[k]] = axB[k];
  const HT Bt = transpose(B, perm);
  double s = 0.0;
  for (int64_t i = 0; i < A.size(); ++i) s += A.v[i] * Bt.v[i];
  return s;
}

HT add(const HT &a, const HT &b, double alpha, double beta) {
  TN_REQUIRE(a.shape == b.shape, "host add: shapes differ");
  HT c = a;
  for (int64_t i = 0; i < c.size(); ++i) c.v[i] = alpha * a.v[i] + beta * b.v[i];
  return c;
}

void scale_axis(HT &t, int axis, const std::vector<double> &s) {
  const int64_t n = t.shape[axis];
  TN_REQUIRE((int64_t)s.size() >= n, "host scale_axis: too few factors");
  const int64_t inner = prod(t.shape, axis + 1);
  for (int64_t i = 0; i < t.size(); ++i) t.v[i] *= s[(i / inner) % n];
}

HT take(const HT &t, int axis, int64_t n) {
  std::vector<int64_t> ns = t.shape;
  TN_REQUIRE(n <= ns[axis], "host take: range exceeds the leg");
  ns[axis] = n;
  HT out(ns);
  const int64_t inner = prod(t.shape, axis + 1), outer = prod(t.shape, 0, axis);
  for (int64_t o = 0; o < outer; ++o)
    for (int64_t a = 0; a < n; ++a)
      for (int64_t i = 0; i < inner; ++i)
        out.v[(o * n + a) * inner + i] = t.v[(o * t.shape[axis] + a) * inner + i];
  return out;
}

void qr(const std::vector<double> &Ain, int m, int n, std::vector<double> &Q,
        std::vector<double> &R) {
  const int k = std::min(m, n);
  std::vector<double> A = Ain;
  std::vector<double> tau(k, 0.0);
  for (int j = 0; j < k; ++j) {
    // Householder vector of column j below the diagonal: H = I - tau v v^T,
    // v[0] = 1, H x = beta e1 with beta = -sign(alpha) |x|
    const double alpha = A[(size_t)j * n + j];
    double xn = 0.0;
    for (int i = j + 1; i < m; ++i) xn = std::hypot(xn, A[(size_t)i * n + j]);
    if (xn == 0.0) continue;  // tau = 0: H = I
    const double beta = -std::copysign(std::hypot(alpha, xn), alpha);
    tau[j] = (beta - alpha) / beta;
    const double sc = 1.0 / (alpha - beta);
    for (int i = j + 1; i < m; ++i) A[(size_t)i * n + j] *= sc;
    A[(size_t)j * n + j] = beta;
    for (int c = j + 1; c < n; ++c) {
      double w = A[(size_t)j * n + c];
      for (int i = j + 1; i < m; ++i) w += A[(size_t)i * n + j] * A[(size_t)i * n + c];
      w *= tau[j];
      A[(size_t)j * n + c] -= w;
      for (int i = j + 1; i < m; ++i) A[(size_t)i * n + c] -= w * A[(size_t)i * n + j];
    }
  }
  R.assign((size_t)k * n, 0.0);
  for (int i = 0; i < k; ++i)
    for (int c = i; c < n; ++c) R[(size_t)i * n + c] = A[(size_t)i * n + c];
  // Q = H_0 ... H_{k-1} [I_k; 0]
  Q.assign((size_t)m * k, 0.0);
  for (int i = 0; i < k; ++i) Q[(size_t)i * k + i] = 1.0;
  for (int j = k - 1; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    for (int c = 0; c < k; ++c) {
      double w = Q[(size_t)j * k + c];
      for (int i = j + 1; i < m; ++i) w += A[(size_t)i * n + j] * Q[(size_t)i * k + c];
      w *= tau[j];
      Q[(size_t)j * k + c] -= w;
      for (int i = j + 1; i < m; ++i) Q[(size_t)i * k + c] -= w * A[(size_t)i * n + j];
    }
  }
}

void eigh(const std::vector<double> &Ain, int n, std::vector<double> &w,
          std::vector<double> &Z) {
  std::vector<double> A = Ain, V((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
  double frob = 0.0;
  for (double x : A) frob += x * x;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a(i, j) * a(i, j);
    if (!(off > 1e-32 * frob)) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  w.resize(n);
  Z.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    w[j] = a(ord[j], ord[j]);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + j] = V[(size_t)i * n + ord[j]];
  }
}

namespace {

// One-sided Jacobi on the columns of the m x n (m >= n) matrix A.
void svd_tall(const std::vector<double> &Ain, int m, int n, std::vector<double> &U,
              std::vector<double> &s, std::vector<double> &VT) {
  std::vector<double> W((size_t)n * m), V((size_t)n * n, 0.0);  // column-major
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) W[(size_t)j * m + i] = Ain[(size_t)i * n + j];
  for (int j = 0; j < n; ++j) V[(size_t)j * n + j] = 1.0;
  const double eps = 2.220446049250313e-16;
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double al = 0.0, be = 0.0, ga = 0.0;
        const double *wp = &W[(size_t)p * m], *wq = &W[(size_t)q * m];
        for (int i = 0; i < m; ++i) {
          al += wp[i] * wp[i];
          be += wq[i] * wq[i];
          ga += wp[i] * wq[i];
        }
        if (ga == 0.0 || !(std::fabs(ga) > eps * std::sqrt(al * be))) continue;
        rotated = true;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                         (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        double *xp = &W[(size_t)p * m], *xq = &W[(size_t)q * m];
        for (int i = 0; i < m; ++i) {
          const double a = xp[i], b = xq[i];
          xp[i] = c * a - sn * b;
          xq[i] = sn * a + c * b;
        }
        double *vp = &V[(size_t)p * n], *vq = &V[(size_t)q * n];
        for (int i = 0; i < n; ++i) {
          const double a = vp[i], b = vq[i];
          vp[i] = c * a - sn * b;
          vq[i] = sn * a + c * b;
        }
      }
    if (!rotated) break;
  }
  std::vector<double> nrm(n);
  for (int j = 0; j < n; ++j) {
    double x = 0.0;
    for (int i = 0; i < m; ++i) x += W[(size_t)j * m + i] * W[(size_t)j * m + i];
    nrm[j] = std::sqrt(x);
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nrm[x] > nrm[y]; });
  s.resize(n);
  U.assign((size_t)m * n, 0.0);
  VT.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    const int c = ord[j];
    s[j] = nrm[c];
    if (nrm[c] > 0.0)
      for (int i = 0; i < m; ++i) U[(size_t)i * n + j] = W[(size_t)c * m + i] / nrm[c];
    for (int i = 0; i < n; ++i) VT[(size_t)j * n + i] = V[(size_t)c * n + i];
  }
}

}  // namespace

void svd(const std::vector<double> &A, int m, int n, std::vector<double> &U,
         std::vector<double> &s, std::vector<double> &VT) {
  if (m >= n) {
    svd_tall(A, m, n, U, s, VT);
    return;
  }
  // A^T = U' s V'^T  =>  A = V' s U'^T
  std::vector<double> At((size_t)n * m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) At[(size_t)j * m + i] = A[(size_t)i * n + j];
  std::vector<double> U2, VT2;
  svd_tall(At, n, m, U2, s, VT2);  // U2: n x m, VT2: m x m
  U.assign((size_t)m * m, 0.0);
  VT.assign((size_t)m * n, 0.0);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) U[(size_t)i * m + j] = VT2[(size_t)j * m + i];
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < n; ++i) VT[(size_t)j * n + i] = U2[(size_t)i * m + j];
}

void qr(const HT &A, const std::vector<int> &row_legs,
        const std::vector<int> &col_legs, HT &Q, HT &R) {
  std::vector<int> perm = row_legs;
  perm.insert(perm.end(), col_legs.begin(), col_legs.end());
  const HT B = transpose(A, perm);
  std::vector<int64_t> rs, cs;
  for (int i : row_legs) rs.push_back(A.shape[i]);
  for (int i : col_legs) cs.push_back(A.shape[i]);
  const int m = (int)prod(rs), n = (int)prod(cs), k = std::min(m, n);
  std::vector<double> q, r;
  qr(B.v, m, n, q, r);
  rs.push_back(k);
  cs.insert(cs.begin(), k);
  Q.shape = rs;
  Q.v = std::move(q);
  R.shape = cs;
  R.v = std::move(r);
}

void svd(const HT &A, const std::vector<int> &row_legs,
         const std::vector<int> &col_legs, HT &U, std::vector<double> &s,
         HT &VT) {
  std::vector<int> perm = row_legs;
  perm.insert(perm.end(), col_legs.begin(), col_legs.end());
  const HT B = transpose(A, perm);
  std::vector<int64_t> rs, cs;
  for (int i : row_legs) rs.push_back(A.shape[i]);
  for (int i : col_legs) cs.push_back(A.shape[i]);
  const int m = (int)prod(rs), n = (int)prod(cs), k = std::min(m, n);
  std::vector<double> u, vt;
  svd(B.v, m, n, u, s, vt);
  rs.push_back(k);
  cs.insert(cs.begin(), k);
  U.shape = rs;
  U.v = std::move(u);
  VT.shape = cs;
  VT.v = std::move(vt);
}

}  // namespace host
}  // namespace tn
