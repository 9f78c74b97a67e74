// Split-precision gather-GEMM for gfx950: C = alpha * A B + beta * C with
// FP64 operands and result in HBM, computed on the bf16 MFMA
// (v_mfma_f32_16x16x32_bf16) as three products of bf16 halves:
//   x = x_hi + x_lo,  x_hi = bf16(x),  x_lo = bf16(x - x_hi),
//   A B ~= A_hi B_hi + A_hi B_lo + A_lo B_hi       (FP32 accumulation).
// Each operand keeps ~16 significant bits (x - x_hi - x_lo ~ 2^-17 |x|) and
// the dropped A_lo B_lo term is ~2^-16 of a product, so a product carries a
// relative error of order 1e-5 before FP32 accumulation; the matrix rate is
// that of bf16 (2.5 PFLOP/s dense) divided by three.
//
// Addressing is the one of the FP64 gather-GEMM (offset tables or affine
// strides per operand, permutations fused into the loads and the store).
// Tiling: 256 threads = 4 waves (2 x 2), block tile BM x BM x 32, each wave
// (BM/2) x (BM/2) = (BM/32)^2 MFMA 16x16 tiles; the K tile is split into hi
// and lo bf16 halves when staged into LDS, stored [m][k] (k contiguous) so
// that a lane reads its 8 consecutive k of one row as one 16-byte load.
#include "common.hpp"
#include "kernels.hpp"

namespace tn {
namespace kern {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int SBK = 32;   // K tile
constexpr int SPAD = 8;   // row padding (bf16 elements) of the LDS tiles
constexpr int SLD = SBK + SPAD;

// Round-to-nearest-even bf16 bits of a finite float.
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short b) {
  return __uint_as_float((unsigned)b << 16);
}

struct SArgs {
  GemmArgs g;
  int k_per_split;
  double *partial;
};

template <int BM, bool A_MFAST, bool B_NFAST>
__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(SArgs ka) {
  const GemmArgs &g = ka.g;
  constexpr int BN = BM;
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int A_PER = BM * SBK / 256, B_PER = BN * SBK / 256;
  // [buffer][hi/lo][row][k]
  __shared__ __attribute__((aligned(16))) unsigned short As[2][2][BM][SLD];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][2][BN][SLD];

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int bn = blockIdx.x, bm = blockIdx.y;
  const int m0 = bm * BM, n0 = bn * BN;
  const int kbeg = blockIdx.z * ka.k_per_split;
  const int kend = min(g.K, kbeg + ka.k_per_split);
  const int M = g.M, N = g.N;

  // loader: element p of this thread in the A tile (BM x SBK) is
  // e = t + 256 p, lanes walking m (MFAST) or k
  int a_ml[A_PER], a_kl[A_PER], a_moff[A_PER];
  bool a_mok[A_PER];
#pragma unroll
  for (int p = 0; p < A_PER; ++p) {
    const int e = t + 256 * p;
    a_ml[p] = A_MFAST ? e % BM : e / SBK;
    a_kl[p] = A_MFAST ? e / BM : e % SBK;
    const int m = m0 + a_ml[p];
    a_mok[p] = m < M;
    a_moff[p] = a_mok[p] ? (g.a_m_stride ? m * g.a_m_stride : g.a_m[m]) : 0;
  }
  int b_nl[B_PER], b_kl[B_PER], b_noff[B_PER];
  bool b_nok[B_PER];
#pragma unroll
  for (int p = 0; p < B_PER; ++p) {
    const int e = t + 256 * p;
    b_nl[p] = B_NFAST ? e % BN : e / SBK;
    b_kl[p] = B_NFAST ? e / BN : e % SBK;
    const int n = n0 + b_nl[p];
    b_nok[p] = n < N;
    b_noff[p] = b_nok[p] ? (g.b_n_stride ? n * g.b_n_stride : g.b_n[n]) : 0;
  }
  double ra[A_PER], rb[B_PER];
  auto load_tile = [&](int k0) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int kk = k0 + a_kl[p];
      const bool ok = a_mok[p] && kk < kend;
      const int koff = ok ? (g.a_k_stride ? kk * g.a_k_stride : g.a_k[kk]) : 0;
      const double v = g.A[ok ? a_moff[p] + koff : 0];
      ra[p] = ok ? v : 0.0;
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int kk = k0 + b_kl[p];
      const bool ok = b_nok[p] && kk < kend;
      const int koff = ok ? (g.b_k_stride ? kk * g.b_k_stride : g.b_k[kk]) : 0;
      const double v = g.B[ok ? b_noff[p] + koff : 0];
      rb[p] = ok ? v : 0.0;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const float f = (float)ra[p];
      const unsigned short hi = bf16_bits(f);
      As[buf][0][a_ml[p]][a_kl[p]] = hi;
      As[buf][1][a_ml[p]][a_kl[p]] = bf16_bits(f - bf16_float(hi));
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const float f = (float)rb[p];
      const unsigned short hi = bf16_bits(f);
      Bs[buf][0][b_nl[p]][b_kl[p]] = hi;
      Bs[buf][1][b_nl[p]][b_kl[p]] = bf16_bits(f - bf16_float(hi));
    }
  };

  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int li = lane & 15;        // row (A) / column (B) of the 16x16 block
  const int kg = (lane >> 4) * 8;  // first of the lane's 8 k
  const int wrow = wm * (BM / 2), wcol = wn * (BN / 2);

  if (kbeg < kend) {
    load_tile(kbeg);
    store_tile(0);
  }
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += SBK) {
    const bool next = k0 + SBK < kend;
    if (next) load_tile(k0 + SBK);
    bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      ah[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][0][wrow + i * 16 + li][kg]);
      al[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][1][wrow + i * 16 + li][kg]);
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      bh[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][0][wcol + j * 16 + li][kg]);
      bl[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][1][wcol + j * 16 + li][kg]);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    if (next) store_tile(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // output layout of the f32 16x16 MFMAs: lane l, register r -> row
  // 4 (l/16) + r, column l % 16
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wcol + j * 16 + li;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wrow + i * 16 + 4 * (lane >> 4) + r;
        if (m >= M) continue;
        const double v = (double)acc[i][j][r];
        if (ka.partial) {
          ka.partial[((int64_t)blockIdx.z * M + m) * N + n] = v;
        } else {
          double *dst = g.C + g.c_m[m] + g.c_n[n];
          *dst = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * (*dst);
        }
      }
    }
}

template <int BM>
void launch_split(const SArgs &ka, dim3 grid, hipStream_t s) {
  const bool am = ka.g.a_mfast, bnf = ka.g.b_nfast;
  if (am && bnf)
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, true, true>), grid, dim3(256), 0, s, ka);
  else if (am)
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, true, false>), grid, dim3(256), 0, s, ka);
  else if (bnf)
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, false, true>), grid, dim3(256), 0, s, ka);
  else
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, false, false>), grid, dim3(256), 0, s, ka);
}

}  // namespace

const char *gemm_bf16x3(const GemmArgs &g, int bm, int splits, int k_per_split,
                        void *workspace, size_t ws_bytes, hipStream_t s) {
  SArgs ka;
  ka.g = g;
  ka.partial = nullptr;
  ka.k_per_split = k_per_split > 0 ? k_per_split : SBK;
  if (splits > 1) {
    const size_t need = (size_t)splits * g.M * g.N * sizeof(double);
    if (workspace == nullptr || ws_bytes < need) {
      splits = 1;
      ka.k_per_split = g.K > 0 ? g.K : SBK;
    } else {
      ka.partial = static_cast<double *>(workspace);
    }
  }
  if (g.K == 0) ka.k_per_split = 0;
  dim3 grid(ceil_div(g.N, bm), ceil_div(g.M, bm), splits);
  if (bm == 128)
    launch_split<128>(ka, grid, s);
  else
    launch_split<64>(ka, grid, s);
  TN_HIP_CHECK(hipGetLastError());
  if (ka.partial)
    splitk_reduce(ka.partial, splits, g.M, g.N, g.alpha, g.beta, g.C, g.c_m, g.c_n, s);
  return bm == 128 ? "gemm_bf16x3_kernel<128,128>" : "gemm_bf16x3_kernel<64,64>";
}

}  // namespace kern
}  // namespace tn
