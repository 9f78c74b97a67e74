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
// D >= 4 (site_mfma_kernel): the NB blocks of a workgroup are stacked so
// that both products are single matrix products on v_mfma_f64_16x16x4f64,
//   MID = [X_1; ...; X_NB] (NB D^2 x D^2) . TB (D^2 x D^2 P)
//   OUT = TK^T (D^2 x D^2 P) . [M_1 | ... | M_NB] (D^2 P x NB D^2),
// M_q the block's intermediate with rows (kA, kB, p) and columns (boA, boB),
// written by the first product straight into LDS in that layout.  At D = 5
// the tiles are padded to 112 x 64 x 28 and 32 x 112 x 52 (62% and 67% of
// the MFMA work is useful).  D <= 3 (site_apply_kernel): the small products
// run on the FP64 VALU with one operand column per thread held in registers.
// Sizes are compile-time constants.
#include "common.hpp"
#include "kernels.hpp"

namespace tn {
namespace kern {

namespace {

constexpr int kSiteThreads = 256;
constexpr int kSiteNB = 4;

// Store the NB staged output blocks os[q][(oA,oB)][(boA,boB)] of workgroup
// (s, f0) in the edge (block-major) or corner layout.
template <bool EDGE, int D, int NT = kSiteThreads>
__device__ __forceinline__ void store_out(const SiteArgs &a, const double *os, int s, int f0,
                                          int nb, int t) {
  constexpr int D2 = D * D;
  constexpr int KO = D2, BO = D2;
  constexpr int NB = kSiteNB;
  if (EDGE) {
    double *dst = a.out + ((int64_t)s * a.Df + f0) * (KO * BO);
    const int n = nb * KO * BO;
    for (int e = t; e < n; e += NT) dst[e] = os[e];
  } else {
    // order (oA, boA, oB, q, boB): runs of nb D contiguous doubles
    const int n = D * D * D * NB * D;
    for (int e = t; e < n; e += NT) {
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

  store_out<EDGE, D>(a, os, s, f0, nb, t);
}

typedef double site_d4 __attribute__((ext_vector_type(4)));

constexpr int kSiteMfmaThreads = 512;  // 8 waves: 4 per SIMD at two workgroups per CU

template <bool EDGE, int D, int P>
__global__ __launch_bounds__(kSiteMfmaThreads) __attribute__((amdgpu_waves_per_eu(4))) void site_mfma_kernel(SiteArgs a) {
  constexpr int D2 = D * D;
  constexpr int KI = D2, BI = D2, KO = D2, BO = D2;
  constexpr int NCB = BO * P;  // intermediate columns (boA, boB, p)
  constexpr int NB = kSiteNB;
  // product 1: M1 x N1 x K1 = NB KI x NCB x BI; product 2: KO x NB BO x KI P
  constexpr int MT1 = (NB * KI + 15) / 16, NT1 = (NCB + 15) / 16, KS1 = (BI + 3) / 4;
  constexpr int MT2 = (KO + 15) / 16, NT2 = (NB * BO + 15) / 16, KS2 = (KI * P + 3) / 4;
  // LDS row strides, odd so that the transposing stores of X (consecutive
  // lanes walk bi, i.e. rows of A1s) and of the intermediate spread over the
  // banks: at 112 doubles the lanes of a row group met on two bank offsets
  constexpr int LA1 = MT1 * 16 + 1;  // row length of A1s (m)
  constexpr int LB2 = NT2 * 16 + 1;  // row length of B2s (n)
  constexpr int NTH = kSiteMfmaThreads;
  constexpr int NW = NTH / 64;
  // product 1: output tiles tau = w, w + NW, ...; tile tau is (mt, nt) =
  // (tau / NT1, tau % NT1), so a wave keeps one N tile (NW % NT1 == 0) and
  // its TB columns in registers
  static_assert(NW % NT1 == 0, "one product-1 N tile per wave");
  constexpr int MTW1 = (MT1 * NT1 + NW - 1) / NW;  // product-1 tiles per wave
  constexpr int NTW2 = (NT2 + NW - 1) / NW;        // product-2 N tiles per wave
  // A1s[k = bi][m = q KI + ki] (zero padded); B2s[k = ki P + p][n = q BO + bo];
  // the output staging os aliases A1s once the first product has read it
  __shared__ double A1s[KS1 * 4 * LA1];
  __shared__ double B2s[KS2 * 4 * LB2];
  static_assert(KS1 * 4 * LA1 >= NB * KO * BO, "output staging fits in A1s");
  double *os = A1s;
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int ngrp = (a.Df + NB - 1) / NB;
  const int s = blockIdx.x / ngrp;
  const int f0 = (blockIdx.x - s * ngrp) * NB;
  const int nb = min(NB, a.Df - f0);

  for (int e = t; e < KS1 * 4 * LA1; e += NTH) A1s[e] = 0.0;
  for (int e = t; e < KS2 * 4 * LB2; e += NTH) B2s[e] = 0.0;
  // operands that stay in registers: this wave's TB columns (product 1) and
  // all of TK^T (product 2)
  const int nt1 = w % NT1;
  double b1[KS1];
  {
    const int c = 16 * nt1 + li;
#pragma unroll
    for (int ks = 0; ks < KS1; ++ks) {
      const int k = 4 * ks + lk;
      b1[ks] = (k < BI && c < NCB) ? a.TB[k * NCB + c] : 0.0;
    }
  }
  double a2[MT2][KS2];
#pragma unroll
  for (int mt = 0; mt < MT2; ++mt) {
    const int m = 16 * mt + li;
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
      const int k = 4 * ks + lk;
      a2[mt][ks] = (k < KI * P && m < KO) ? a.TK[k * KO + m] : 0.0;
    }
  }
  __syncthreads();
  // ---- load: one contiguous run, transposed into A1s ----------------------
  {
    const double *src = a.X + ((int64_t)s * a.Df + f0) * (KI * BI);
    const int n = nb * KI * BI;
    for (int e = t; e < n; e += NTH) {
      const int q = e / (KI * BI), r = e - q * (KI * BI);
      const int ki = r / BI, bi = r - ki * BI;
      A1s[bi * LA1 + q * KI + ki] = src[e];
    }
  }
  __syncthreads();
  // ---- product 1 -> B2s -----------------------------------------------------
  {
    site_d4 acc[MTW1];
#pragma unroll
    for (int j = 0; j < MTW1; ++j) acc[j] = site_d4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS1; ++ks) {
#pragma unroll
      for (int j = 0; j < MTW1; ++j) {
        const int mt = (w + NW * j) / NT1;
        if (mt < MT1) {
          const double av = A1s[(4 * ks + lk) * LA1 + 16 * mt + li];
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b1[ks], acc[j], 0, 0, 0);
        }
      }
    }
    // lane l, register r: row 16 mt + l/16 + 4 r, column 16 nt1 + l%16
    const int c = 16 * nt1 + li;
    const int p = c % P, bo = c / P;
#pragma unroll
    for (int j = 0; j < MTW1; ++j) {
      const int mt = (w + NW * j) / NT1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 16 * mt + lk + 4 * r;
        if (mt < MT1 && m < NB * KI && c < NCB) {
          const int q = m / KI, ki = m - q * KI;
          B2s[(ki * P + p) * LB2 + q * BO + bo] = acc[j][r];
        }
      }
    }
  }
  __syncthreads();
  // ---- product 2 -> os (over A1s, free since the barrier) -------------------
#pragma unroll
  for (int j = 0; j < NTW2; ++j) {
    const int nt = w + NW * j;
    if (nt < NT2) {
      site_d4 acc[MT2];
#pragma unroll
      for (int mt = 0; mt < MT2; ++mt) acc[mt] = site_d4{0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KS2; ++ks) {
        const double bv = B2s[(4 * ks + lk) * LB2 + 16 * nt + li];
#pragma unroll
        for (int mt = 0; mt < MT2; ++mt)
          acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[mt][ks], bv, acc[mt], 0, 0, 0);
      }
      const int n = 16 * nt + li;
#pragma unroll
      for (int mt = 0; mt < MT2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = 16 * mt + lk + 4 * r;
          if (co < KO && n < NB * BO) {
            const int q = n / BO, bo = n - q * BO;
            os[q * KO * BO + co * BO + bo] = acc[mt][r];
          }
        }
    }
  }
  __syncthreads();
  store_out<EDGE, D, NTH>(a, os, s, f0, nb, t);
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
      hipLaunchKernelGGL((site_mfma_kernel<EDGE, 4, 2>), grid, dim3(kSiteMfmaThreads), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((site_mfma_kernel<EDGE, 5, 2>), grid, dim3(kSiteMfmaThreads), 0, s, a);
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
