// Transfer-matrix spectrum of the CTM environment on the GPU.  The chi^2-long
// Krylov vectors, the matvecs and the Gram-Schmidt projections live on the
// device; only the (maxvec+1) x maxvec Hessenberg matrix is assembled on the
// host, where its eigenvalues are computed (TeNeS does the same with dgeev).
#include "transfer.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "eig_host.hpp"
#include "kernels.hpp"
#include "small_dense.hpp"

namespace tenes {
namespace itps {

using tn::Axes;
using tn::Shape;

// transfer_matrix.cpp:503-525
size_t TransferMatrix_ctm::dim(int dir, int fixed_coord) const {
  if (dir == 0) {
    const int s0 = lattice.index(0, fixed_coord);
    const int s1 = lattice.top(s0);
    return (size_t)(eTt[s0].shape[0] * eTb[s1].shape[1]);
  }
  const int s0 = lattice.index(fixed_coord, 0);
  const int s1 = lattice.left(s0);
  return (size_t)(eTl[s0].shape[0] * eTr[s1].shape[1]);
}

// transfer_matrix.cpp:162-185
tensor TransferMatrix_ctm::initial_vector(int dir, int fixed_coord) const {
  tensor v;
  if (dir == 0) {
    const int site = lattice.index(0, fixed_coord);
    v = tn::tensordot(C1[site], C4[lattice.top(site)], Axes{0}, Axes{1});
  } else {
    const int site = lattice.index(fixed_coord, 0);
    v = tn::tensordot(C4[site], C3[lattice.left(site)], Axes{0}, Axes{1});
  }
  return tn::reshape(v, Shape{v.size()});
}

// transfer_matrix.cpp:187-308 (double-layer branch)
tensor TransferMatrix_ctm::matvec(int dir, const tensor &invec,
                                  int fixed_coord) const {
  if (dir == 0) {
    const int y = fixed_coord;
    const int s0 = lattice.index(0, y);
    const int s1 = lattice.top(s0);
    const int64_t CHI0 = eTt[s0].shape[0], CHI1 = eTb[s1].shape[1];
    tensor out = tn::reshape(invec, Shape{CHI0, CHI1});
    for (int x = 0; x < lattice.LX; ++x) {
      const int site = lattice.index(x, y);
      const tensor &top = eTt[site];
      const tensor &bottom = eTb[lattice.top(site)];
      out = tn::tensordot(top, tn::tensordot(bottom, out, Axes{1}, Axes{1}),
                          Axes{0, 2, 3}, Axes{3, 1, 2});
    }
    return tn::reshape(out, Shape{CHI0 * CHI1});
  }
  int x = fixed_coord;
  const int x_orig = x;
  const int s0 = lattice.index(x, 0);
  const int s1 = lattice.left(s0);
  const int64_t CHI0 = eTl[s0].shape[0], CHI1 = eTr[s1].shape[1];
  tensor out = tn::reshape(invec, Shape{CHI0, CHI1});
  do {
    for (int y = 0; y < lattice.LY; ++y) {
      const int site = lattice.index(x, y);
      const tensor &left = eTl[site];
      const tensor &right = eTr[lattice.left(site)];
      out = tn::tensordot(left, tn::tensordot(right, out, Axes{1}, Axes{1}),
                          Axes{0, 2, 3}, Axes{3, 1, 2});
    }
    x = (x + lattice.skew + lattice.LX) % lattice.LX;
  } while (x != x_orig);
  return tn::reshape(out, Shape{CHI0 * CHI1});
}

// transfer_matrix.cpp:379-475 (double-layer branch)
tensor TransferMatrix_ctm::matrix(int dir, int fixed_coord) const {
  tensor res;
  int64_t CHI0, CHI1;
  if (dir == 0) {
    const int y = fixed_coord;
    const int s0 = lattice.index(0, y);
    const int s1 = lattice.top(s0);
    CHI0 = eTt[s0].shape[0];
    CHI1 = eTb[s1].shape[1];
    res = tn::tensordot_t(eTt[s0], eTb[s1], Axes{2, 3}, Axes{2, 3},
                          Axes{0, 3, 1, 2});
    for (int x = 1; x < lattice.LX; ++x) {
      const int site = lattice.index(x, y);
      res = tn::tensordot(
          res,
          tn::tensordot(eTt[site], eTb[lattice.top(site)], Axes{2, 3},
                        Axes{2, 3}),
          Axes{2, 3}, Axes{0, 3});
    }
  } else {
    int x = fixed_coord;
    const int x_orig = x;
    const int s0 = lattice.index(x, 0);
    const int s1 = lattice.left(s0);
    CHI0 = eTl[s0].shape[0];
    CHI1 = eTr[s1].shape[1];
    std::vector<double> id((size_t)(CHI0 * CHI1 * CHI0 * CHI1), 0.0);
    for (int64_t i = 0; i < CHI0 * CHI1; ++i)
      id[(size_t)(i * CHI0 * CHI1 + i)] = 1.0;
    res = tn::from_host(*C1[0].ctx(), Shape{CHI0, CHI1, CHI0, CHI1}, id.data());
    do {
      for (int y = 0; y < lattice.LY; ++y) {
        const int site = lattice.index(x, y);
        res = tn::tensordot(
            res,
            tn::tensordot(eTl[site], eTr[lattice.left(site)], Axes{2, 3},
                          Axes{2, 3}),
            Axes{2, 3}, Axes{0, 3});
      }
      x = (x + lattice.skew + lattice.LX) % lattice.LX;
    } while (x != x_orig);
  }
  return tn::reshape(res, Shape{CHI0 * CHI1, CHI0 * CHI1});
}

// transfer_matrix.cpp:84-142 + arnoldi.cpp:56-181
std::vector<std::complex<double>> TransferMatrix_ctm::eigenvalues(
    int dir, int fixed_coord, const TransferMatrix_Parameters &params) const {
  tn::Context &ctx = *C1[0].ctx();
  const size_t N = dim(dir, fixed_coord);
  const int nev = (int)std::min<size_t>((size_t)params.num_eigvals, N);
  std::vector<std::complex<double>> eig;
  if (N == 1) {
    eig.assign(1, std::complex<double>(1.0, 0.0));
    return eig;
  }
  if ((int64_t)N <= params.maxdim_dense_eigensolver) {
    tensor M = matrix(dir, fixed_coord);
    std::vector<double> h(N * N);
    tn::to_host(M, h.data());
    tn::eig_nonsym((int)N, h, nev, eig, nullptr);
    return eig;
  }
  int maxvec = params.arnoldi_maxdim;
  int maxiter = params.arnoldi_maxiter;
  if ((int64_t)N < maxvec) {
    maxvec = (int)N;
    maxiter = 1;  // transfer_matrix.cpp:117-120
  }
  // arnoldi.cpp:60-67 writes Q[1..nev] before it looks at maxvec: more
  // eigenvalues than basis vectors would run past the basis
  if (nev > maxvec)
    throw tn::Error("transfer matrix: num_eigvals (" + std::to_string(nev) +
                        ") exceeds arnoldi_maxdim (" + std::to_string(maxvec) + ")",
                    -3);
  const int mindim = std::max(params.arnoldi_restartdim, nev);  // arnoldi.cpp:63-65
  hipStream_t s = ctx.stream();
  tensor Qall = tn::empty(ctx, Shape{maxvec + 1, (int64_t)N});
  auto Qp = [&](int k) { return Qall.data() + (int64_t)k * N; };
  tensor Hdev = tn::zeros(ctx, Shape{maxvec + 1, maxvec});
  std::vector<double> Hh((size_t)(maxvec + 1) * maxvec, 0.0);
  const int np = tn::kern::dot_partials((int64_t)N);
  tensor part = tn::empty(ctx, Shape{np});
  tensor nrm2 = tn::empty(ctx, Shape{1});

  // arnoldi.cpp:98-112: modified Gram-Schmidt of Q[k] against Q[0..k-1];
  // H[j][k-1] = <Q[j], Q[k]>, H[k][k-1] = ||Q[k]||.
  auto orthonormalize = [&](int k) {
    for (int j = 0; j < k; ++j) {
      double *hjk = Hdev.data() + (int64_t)j * maxvec + (k - 1);
      tn::kern::dot(Qp(j), Qp(k), (int64_t)N, hjk, part.data(), s);
      tn::kern::axpy_dev(Qp(k), Qp(j), (int64_t)N, hjk, -1.0, s);
    }
    tn::kern::dot(Qp(k), Qp(k), (int64_t)N, nrm2.data(), part.data(), s);
    double n2 = 0.0;
    TN_HIP_CHECK(hipMemcpyAsync(&n2, nrm2.data(), sizeof(double),
                                hipMemcpyDeviceToHost, s));
    if (k > 0) {
      for (int j = 0; j < k; ++j)
        TN_HIP_CHECK(hipMemcpyAsync(&Hh[(size_t)j * maxvec + (k - 1)],
                                    Hdev.data() + (int64_t)j * maxvec + (k - 1),
                                    sizeof(double), hipMemcpyDeviceToHost, s));
    }
    ctx.sync();
    const double x = std::sqrt(n2);
    tn::kern::scale(Qp(k), (int64_t)N, 1.0 / x, s);
    if (k > 0) Hh[(size_t)k * maxvec + (k - 1)] = x;
  };
  auto apply = [&](int k) {
    tensor in = tn::borrow(ctx, Shape{(int64_t)N}, Qp(k - 1));
    tensor out = matvec(dir, in, fixed_coord);
    TN_HIP_CHECK(hipMemcpyAsync(Qp(k), out.data(), N * sizeof(double),
                                hipMemcpyDeviceToDevice, s));
  };
  // arnoldi.cpp:169-181
  auto residue = [&](int k) {
    std::vector<double> h((size_t)k * k);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) h[(size_t)i * k + j] = Hh[(size_t)i * maxvec + j];
    std::vector<std::complex<double>> ev, last;
    tn::eig_nonsym(k, h, nev, ev, &last);
    const double hk = Hh[(size_t)k * maxvec + (k - 1)];
    std::vector<double> res(nev);
    for (int i = 0; i < nev; ++i) res[i] = std::abs(hk * last[i]) / std::abs(ev[i]);
    return res;
  };

  // arnoldi.cpp:117-167: maxvec - minvec QR steps h <- q^T h q of the
  // Hessenberg block, applied to the basis, then entries of the block below
  // 1e-12 in modulus zeroed.  Each reference step factors
  // h - identity(k, sigma) with sigma a Ritz value, but identity
  // (tensor.hpp:58-64) sets the out-of-range element (k, k) in place of
  // (i, i) and returns a zero matrix: every step is an unshifted QR step (the
  // new start vector is A Q[0] / r_00, a power step), and so it is here.  The
  // product of the steps' q's turns the basis once, Q[j] <- sum_i Q[i] qt[i][j]
  // for j < maxvec, as a (maxvec x maxvec) x (maxvec x N) product on the
  // device.  Later steps overwrite only what orthonormalize writes; the rest
  // of the restarted H stays in place, as in the single H array of
  // arnoldi.cpp.
  auto restart = [&](int minvec) {
    const int k = maxvec;
    std::vector<double> h((size_t)k * k);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) h[(size_t)i * k + j] = Hh[(size_t)i * maxvec + j];
    std::vector<double> qt((size_t)k * k, 0.0), q, r, t((size_t)k * k);
    for (int i = 0; i < k; ++i) qt[(size_t)i * k + i] = 1.0;
    for (int step = 0; step < k - minvec; ++step) {
      tn::host::qr(h, k, k, q, r);
      // h <- q^T (h q)
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
          double acc = 0.0;
          for (int l = 0; l < k; ++l) acc += h[(size_t)i * k + l] * q[(size_t)l * k + j];
          t[(size_t)i * k + j] = acc;
        }
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
          double acc = 0.0;
          for (int l = 0; l < k; ++l) acc += q[(size_t)l * k + i] * t[(size_t)l * k + j];
          h[(size_t)i * k + j] = acc;
        }
      // qt <- qt q
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
          double acc = 0.0;
          for (int l = 0; l < k; ++l) acc += qt[(size_t)i * k + l] * q[(size_t)l * k + j];
          t[(size_t)i * k + j] = acc;
        }
      qt.swap(t);
    }
    if (k > minvec) {
      // new Q[j] = sum_i qt[i][j] Q[i]: (qt^T) x Q
      std::vector<double> qtt((size_t)k * k);
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) qtt[(size_t)j * k + i] = qt[(size_t)i * k + j];
      tensor Qt = tn::from_host(ctx, Shape{k, k}, qtt.data());
      tensor Qk = tn::borrow(ctx, Shape{k, (int64_t)N}, Qall.data());
      tensor Qn = tn::tensordot(Qt, Qk, Axes{1}, Axes{0});
      TN_HIP_CHECK(hipMemcpyAsync(Qall.data(), Qn.data(), (size_t)k * N * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
    }
    const double cut2 = 1.0e-12 * 1.0e-12;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        double v = h[(size_t)i * k + j];
        Hh[(size_t)i * maxvec + j] = v * v < cut2 ? 0.0 : v;
      }
  };

  tensor v0 = initial_vector(dir, fixed_coord);
  TN_HIP_CHECK(hipMemcpyAsync(Qp(0), v0.data(), N * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
  orthonormalize(0);
  for (int k = 1; k <= nev; ++k) {
    apply(k);
    orthonormalize(k);
  }
  // arnoldi.cpp:68-84
  double res_nev = 1.0;
  for (int iter = 1; iter <= maxiter; ++iter) {
    for (int k = nev + 1; k <= maxvec; ++k) {
      apply(k);
      orthonormalize(k);
      res_nev = residue(k)[nev - 1];
      if (res_nev <= params.arnoldi_rtol) break;
    }
    if (res_nev <= params.arnoldi_rtol || iter == maxiter) break;
    restart(mindim);
  }
  std::vector<double> h((size_t)maxvec * maxvec);
  for (int i = 0; i < maxvec; ++i)
    for (int j = 0; j < maxvec; ++j)
      h[(size_t)i * maxvec + j] = Hh[(size_t)i * maxvec + j];
  tn::eig_nonsym(maxvec, h, nev, eig, nullptr);
  return eig;
}

}  // namespace itps
}  // namespace tenes
