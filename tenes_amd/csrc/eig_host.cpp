// Host eigen-solver for small nonsymmetric matrices:
//   Householder reduction to upper Hessenberg form, then the complex
//   single-shift QR algorithm (Wilkinson shifts, Givens rotations, standard
//   small-subdiagonal deflation), then inverse iteration for the few right
//   eigenvectors whose last component the Arnoldi residual needs.
#include "eig_host.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

#include "common.hpp"

namespace tn {

namespace {

using cd = std::complex<double>;

void to_hessenberg(int n, std::vector<double> &A) {
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += A[i * n + k] * A[i * n + k];
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    std::vector<double> v(n, 0.0);
    const double x0 = A[(k + 1) * n + k];
    const double sgn = x0 >= 0.0 ? 1.0 : -1.0;
    for (int i = k + 1; i < n; ++i) v[i] = A[i * n + k];
    v[k + 1] += sgn * alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    // A <- (I - 2 v v^T / vn) A (I - 2 v v^T / vn)
    for (int j = 0; j < n; ++j) {
      double d = 0.0;
      for (int i = k + 1; i < n; ++i) d += v[i] * A[i * n + j];
      d *= 2.0 / vn;
      for (int i = k + 1; i < n; ++i) A[i * n + j] -= d * v[i];
    }
    for (int i = 0; i < n; ++i) {
      double d = 0.0;
      for (int j = k + 1; j < n; ++j) d += A[i * n + j] * v[j];
      d *= 2.0 / vn;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= d * v[j];
    }
    for (int i = k + 2; i < n; ++i) A[i * n + k] = 0.0;
  }
}

void hessenberg_qr(int n, std::vector<cd> &H, std::vector<cd> &ev) {
  ev.assign(n, cd(0.0, 0.0));
  double hnorm = 0.0;
  for (auto &v : H) hnorm = std::max(hnorm, std::abs(v));
  int hi = n - 1;
  int iter = 0;
  while (hi >= 0) {
    if (hi == 0) {
      ev[0] = H[0];
      break;
    }
    int lo = hi;
    while (lo > 0) {
      double s = std::abs(H[lo * n + lo]) + std::abs(H[(lo - 1) * n + lo - 1]);
      if (s == 0.0) s = hnorm;
      if (std::abs(H[lo * n + lo - 1]) <= DBL_EPSILON * s) {
        H[lo * n + lo - 1] = 0.0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      ev[hi] = H[hi * n + hi];
      --hi;
      iter = 0;
      continue;
    }
    ++iter;
    TN_REQUIRE(iter < 300, "eigenvalue QR iteration did not converge");
    cd mu;
    if (iter % 11 == 10) {
      mu = H[hi * n + hi] + 0.75 * std::abs(H[hi * n + hi - 1]);
    } else {
      const cd a = H[(hi - 1) * n + hi - 1], b = H[(hi - 1) * n + hi];
      const cd c = H[hi * n + hi - 1], d = H[hi * n + hi];
      const cd half = 0.5 * (a - d);
      const cd disc = std::sqrt(half * half + b * c);
      const cd l1 = 0.5 * (a + d) + disc, l2 = 0.5 * (a + d) - disc;
      mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
    }
    for (int k = lo; k <= hi; ++k) H[k * n + k] -= mu;
    std::vector<cd> cs(hi - lo), ss(hi - lo);
    for (int k = lo; k < hi; ++k) {
      const cd a = H[k * n + k], b = H[(k + 1) * n + k];
      const double r = std::hypot(std::abs(a), std::abs(b));
      cd c(1.0, 0.0), s(0.0, 0.0);
      if (r > 0.0) {
        c = a / r;
        s = b / r;
      }
      cs[k - lo] = c;
      ss[k - lo] = s;
      for (int j = k; j <= hi; ++j) {
        const cd x = H[k * n + j], y = H[(k + 1) * n + j];
        H[k * n + j] = std::conj(c) * x + std::conj(s) * y;
        H[(k + 1) * n + j] = -s * x + c * y;
      }
    }
    for (int k = lo; k < hi; ++k) {
      const cd c = cs[k - lo], s = ss[k - lo];
      const int rmax = std::min(k + 2, hi);
      for (int i = lo; i <= rmax; ++i) {
        const cd x = H[i * n + k], y = H[i * n + k + 1];
        H[i * n + k] = x * c + y * s;
        H[i * n + k + 1] = -x * std::conj(s) + y * std::conj(c);
      }
    }
    for (int k = lo; k <= hi; ++k) H[k * n + k] += mu;
  }
}

// Unit-norm right eigenvector of A for eigenvalue lam by inverse iteration;
// returns its last component.
cd last_component(int n, const std::vector<double> &A, cd lam) {
  double anorm = 0.0;
  for (auto v : A) anorm = std::max(anorm, std::fabs(v));
  if (anorm == 0.0) anorm = 1.0;
  const double tiny = DBL_EPSILON * anorm;
  std::vector<cd> M(n * n);
  const cd shift = lam + cd(tiny, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      M[i * n + j] = cd(A[i * n + j], 0.0) - (i == j ? shift : cd(0.0, 0.0));
  std::vector<int> piv(n);
  std::iota(piv.begin(), piv.end(), 0);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(M[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::abs(M[i * n + k]) > best) {
        best = std::abs(M[i * n + k]);
        p = i;
      }
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(M[k * n + j], M[p * n + j]);
      std::swap(piv[k], piv[p]);
    }
    if (std::abs(M[k * n + k]) < tiny) M[k * n + k] = tiny;
    for (int i = k + 1; i < n; ++i) {
      const cd f = M[i * n + k] / M[k * n + k];
      M[i * n + k] = f;
      for (int j = k + 1; j < n; ++j) M[i * n + j] -= f * M[k * n + j];
    }
  }
  std::vector<cd> x(n, cd(1.0, 0.0)), y(n);
  for (int it = 0; it < 3; ++it) {
    for (int i = 0; i < n; ++i) y[i] = x[piv[i]];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) y[i] -= M[i * n + j] * y[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) y[i] -= M[i * n + j] * y[j];
      y[i] /= M[i * n + i];
    }
    double nrm = 0.0;
    for (auto &v : y) nrm += std::norm(v);
    nrm = std::sqrt(nrm);
    for (int i = 0; i < n; ++i) x[i] = y[i] / nrm;
  }
  return x[n - 1];
}

}  // namespace

void eig_nonsym(int n, const std::vector<double> &A, int nev,
                std::vector<std::complex<double>> &eigvals,
                std::vector<std::complex<double>> *last) {
  TN_REQUIRE(n >= 1 && (int)A.size() == n * n, "eig_nonsym: bad size");
  std::vector<double> Hr = A;
  to_hessenberg(n, Hr);
  std::vector<cd> H(n * n);
  for (int i = 0; i < n * n; ++i) H[i] = cd(Hr[i], 0.0);
  std::vector<cd> ev;
  hessenberg_qr(n, H, ev);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    return std::norm(ev[a]) > std::norm(ev[b]);
  });
  nev = std::min(nev, n);
  eigvals.resize(nev);
  for (int i = 0; i < nev; ++i) eigvals[i] = ev[idx[i]];
  if (last) {
    last->resize(nev);
    for (int i = 0; i < nev; ++i) (*last)[i] = last_component(n, A, eigvals[i]);
  }
}

}  // namespace tn
