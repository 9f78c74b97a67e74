// Double-layer site application, the HBM-bound pair of contractions in the
// enlarged corners (ctm.cpp:207-227) and in Calc_Next_eT (ctm.cpp:532-536)
// as one kernel.  For every block (s, f) of the two environment (chi) legs
// of X, with inner legs kA, kB (met by the ket site tensor) and bA, bB (met
// by the bra):
//   mid[(kA,kB), (boA,boB), p] = sum_{bA,bB} X[(kA,kB),(bA,bB)] TB[(bA,bB),(boA,boB),p]
//   out[(oA,oB), (boA,boB)]    = sum_{kA,kB,p} TK[(kA,kB),p,(oA,oB)] mid[(kA,kB),(boA,boB),p]
// The intermediate (D^2 d values per block, 100 MB over the whole tensor at
// D=5, chi=100) stays in LDS, so HBM sees X once and the result once.
//
// Layouts (row-major): X (s, f, kA, kB, bA, bB) block-major (the GEMM that
// produces X writes it so, for free); out either
//   corner: (oA, s, boA, oB, f, boB)   (the enlarged corner's matrix layout)
//   edge:   (s, f, oA, oB, boA, boB)   (block-major).
// A workgroup takes NB consecutive f of one s: its loads are one contiguous
// run of NB D^4 doubles, its stores runs of NB D (corner) or NB D^4 (edge).
// The small products run on the FP64 VALU with one operand column per
// thread held in registers (a 25 x 25 x 50 product would waste most of a
// 16 x 16 x 4 MFMA tile to padding); the sizes are compile-time constants.
#include "common.hpp"
#include "kernels.hpp"

namespace tn {
namespace kern {

namespace {

constexpr int kSiteThreads = 256;
constexpr int kSiteNB = 4;

template <bool EDGE, int D, int P>
__global__ __launch_bounds__(kSiteThreads) void site_apply_kernel(SiteArgs a) {
  constexpr int D2 = D * D;
  constexpr int KI = D2, BI = D2, KO = D2, BO = D2;
  constexpr int NCB = BO * P;               // mid columns (boA, boB, p)
  constexpr int GB = kSiteThreads / NCB;    // row groups of the mid product
  constexpr int RB = (KI + GB - 1) / GB;
  constexpr int GC = kSiteThreads / KO;     // column groups of the output
  constexpr int RC = (BO + GC - 1) / GC;
  constexpr int NB = kSiteNB;
  __shared__ double xs[NB * KI * BI];
  __shared__ double os[NB * KO * BO];
  __shared__ double mid[KI * NCB];
  const int t = threadIdx.x;
  const int ngrp = (a.Df + NB - 1) / NB;
  const int s = blockIdx.x / ngrp;
  const int f0 = (blockIdx.x - s * ngrp) * NB;
  const int nb = min(NB, a.Df - f0);

  // ---- load: one contiguous run -------------------------------------------
  {
    const double *src = a.X + ((int64_t)s * a.Df + f0) * (KI * BI);
    const int n = nb * KI * BI;
    for (int e = t; e < n; e += kSiteThreads) xs[e] = src[e];
  }
  // per-thread operand columns (registers), fixed for all blocks
  const int cb = t % NCB, gb = t / NCB;
  double tb[BI];
#pragma unroll
  for (int i = 0; i < BI; ++i) tb[i] = gb < GB ? a.TB[i * NCB + cb] : 0.0;
  const int co = t % KO, gc = t / KO;
  double tk[KI * P];
#pragma unroll
  for (int i = 0; i < KI * P; ++i) tk[i] = gc < GC ? a.TK[i * KO + co] : 0.0;
  __syncthreads();

  for (int q = 0; q < nb; ++q) {
    const double *xq = xs + q * KI * BI;
    if (gb < GB) {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int ki = gb + GB * r;
        if (ki < KI) {
          double acc = 0.0;
#pragma unroll
          for (int bi = 0; bi < BI; ++bi) acc += xq[ki * BI + bi] * tb[bi];
          mid[ki * NCB + cb] = acc;
        }
      }
    }
    __syncthreads();
    if (gc < GC) {
#pragma unroll
      for (int r = 0; r < RC; ++r) {
        const int bo = gc + GC * r;
        if (bo < BO) {
          double acc = 0.0;
#pragma unroll
          for (int ki = 0; ki < KI; ++ki)
#pragma unroll
            for (int p = 0; p < P; ++p) acc += tk[ki * P + p] * mid[ki * NCB + bo * P + p];
          os[q * KO * BO + co * BO + bo] = acc;
        }
      }
    }
    __syncthreads();
  }

  // ---- store ---------------------------------------------------------------
  if (EDGE) {
    double *dst = a.out + ((int64_t)s * a.Df + f0) * (KO * BO);
    const int n = nb * KO * BO;
    for (int e = t; e < n; e += kSiteThreads) dst[e] = os[e];
  } else {
    // order (oA, boA, oB, q, boB): runs of nb D contiguous doubles
    const int n = D * D * D * NB * D;
    for (int e = t; e < n; e += kSiteThreads) {
      int r = e;
      const int boB = r % D;
      r /= D;
      const int q = r % NB;
      r /= NB;
      const int oB = r % D;
      r /= D;
      const int boA = r % D;
      const int oA = r / D;
      if (q >= nb) continue;
      const int64_t off =
          (((((int64_t)oA * a.Ds + s) * D + boA) * D + oB) * a.Df + f0 + q) * D + boB;
      a.out[off] = os[q * KO * BO + (oA * D + oB) * BO + boA * D + boB];
    }
  }
}

template <bool EDGE>
void launch_site(const SiteArgs &a, dim3 grid, hipStream_t s) {
  switch (a.D) {
    case 2:
      hipLaunchKernelGGL((site_apply_kernel<EDGE, 2, 2>), grid, dim3(kSiteThreads), 0, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((site_apply_kernel<EDGE, 3, 2>), grid, dim3(kSiteThreads), 0, s, a);
      break;
    case 4:
      hipLaunchKernelGGL((site_apply_kernel<EDGE, 4, 2>), grid, dim3(kSiteThreads), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((site_apply_kernel<EDGE, 5, 2>), grid, dim3(kSiteThreads), 0, s, a);
      break;
  }
}

}  // namespace

bool site_apply_supported(int D, int p) { return p == 2 && D >= 2 && D <= 5; }

void site_apply(const SiteArgs &a, bool edge, hipStream_t s) {
  TN_REQUIRE(site_apply_supported(a.D, a.p) && a.Ds >= 1 && a.Df >= 1,
             "site_apply: shape outside the kernel's limits");
  const int ngrp = (a.Df + kSiteNB - 1) / kSiteNB;
  const dim3 grid((unsigned)(a.Ds * ngrp));
  if (edge)
    launch_site<true>(a, grid, s);
  else
    launch_site<false>(a, grid, s);
  TN_HIP_CHECK(hipGetLastError());
}

}  // namespace kern
}  // namespace tn
