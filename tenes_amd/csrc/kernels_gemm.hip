// FP64 gather-GEMM for gfx950 (CDNA4) on v_mfma_f64_16x16x4_f64.
//
// One kernel serves every tensordot of the CTMRG path.  Operands are tensors
// in row-major HBM layout; the tensordot's leg permutation is expressed as
// two offset tables per operand (outer index, contracted index), so the
// permutation is fused into the global->LDS tile loads and the output
// transpose into the epilogue store.  When a table is an arithmetic
// progression the loader computes the offset directly (affine fast path).
//
// Tiling: 256 threads = 4 waves (2 x 2), block tile BM x BN x 16, each wave
// owns (BM/2) x (BN/2) = (BM/32) x (BN/32) MFMA 16x16 accumulator tiles.
// K tiles are staged through LDS with a register prefetch of the next tile.
// An FP64 16x16x4 MFMA does 2048 flop; at the 78.6 TFLOP/s FP64 matrix peak
// (256 CUs x 4 SIMDs) that is ~64 cycles per instruction per SIMD, so the LDS
// traffic per MFMA (one double per lane per operand) is far below the LDS
// bandwidth; the design goal is to keep the MFMA pipe fed, and to have enough
// workgroups (>= 256 CUs) in flight, splitting K when M x N is small.
#include "common.hpp"
#include "kernels.hpp"

namespace tn {
namespace kern {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int BK = 16;

struct KArgs {
  GemmArgs g;
  int k_per_split;
  double *partial;  // split-K partials [z][M][N] or nullptr
};

template <bool FAST_STRIDE>
__device__ __forceinline__ int off_of(const int *tab, int stride, int i) {
  return FAST_STRIDE ? i * stride : tab[i];
}

template <int BM, int BN, bool A_MFAST, bool B_NFAST, bool AFF>
__global__ __launch_bounds__(256) void gemm_f64_kernel(KArgs ka) {
  const GemmArgs &g = ka.g;
  constexpr int TM = BM / 32;
  constexpr int TN = BN / 32;
  constexpr int A_PER = BM * BK / 256;
  constexpr int B_PER = BN * BK / 256;
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * BM;
  const int n0 = blockIdx.x * BN;
  const int kbeg = blockIdx.z * ka.k_per_split;
  const int kend = min(g.K, kbeg + ka.k_per_split);
  const int M = g.M, N = g.N;

  // ---- loader geometry ----------------------------------------------------
  // A tile: BM x BK.  MFAST: lanes walk m; else lanes walk k.
  constexpr int A_KSTEP = A_MFAST ? (256 / BM) : 0;
  constexpr int A_MSTEP = A_MFAST ? 0 : (256 / BK);
  const int a_ml = A_MFAST ? (t % BM) : (t / BK);
  const int a_kl = A_MFAST ? (t / BM) : (t % BK);
  constexpr int B_KSTEP = B_NFAST ? (256 / BN) : 0;
  constexpr int B_NSTEP = B_NFAST ? 0 : (256 / BK);
  const int b_nl = B_NFAST ? (t % BN) : (t / BK);
  const int b_kl = B_NFAST ? (t / BN) : (t % BK);

  // Outer offsets are fixed for the whole K loop: keep them in registers.
  int a_moff[A_MFAST ? 1 : A_PER];
  bool a_mok[A_MFAST ? 1 : A_PER];
#pragma unroll
  for (int p = 0; p < (A_MFAST ? 1 : A_PER); ++p) {
    const int m = m0 + a_ml + p * A_MSTEP;
    a_mok[p] = m < M;
    a_moff[p] = a_mok[p] ? (AFF && g.a_m_stride ? m * g.a_m_stride : g.a_m[m])
                         : 0;
  }
  int b_noff[B_NFAST ? 1 : B_PER];
  bool b_nok[B_NFAST ? 1 : B_PER];
#pragma unroll
  for (int p = 0; p < (B_NFAST ? 1 : B_PER); ++p) {
    const int n = n0 + b_nl + p * B_NSTEP;
    b_nok[p] = n < N;
    b_noff[p] = b_nok[p] ? (AFF && g.b_n_stride ? n * g.b_n_stride : g.b_n[n])
                         : 0;
  }

  double ra[A_PER], rb[B_PER];
  auto load_tile = [&](int k0) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      int m_i = A_MFAST ? 0 : p;
      int kk = k0 + a_kl + (A_MFAST ? p * A_KSTEP : 0);
      bool ok = a_mok[m_i] && kk < kend;
      int koff = 0;
      if (ok) koff = (AFF && g.a_k_stride) ? kk * g.a_k_stride : g.a_k[kk];
      ra[p] = ok ? g.A[a_moff[m_i] + koff] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      int n_i = B_NFAST ? 0 : p;
      int kk = k0 + b_kl + (B_NFAST ? p * B_KSTEP : 0);
      bool ok = b_nok[n_i] && kk < kend;
      int koff = 0;
      if (ok) koff = (AFF && g.b_k_stride) ? kk * g.b_k_stride : g.b_k[kk];
      rb[p] = ok ? g.B[b_noff[n_i] + koff] : 0.0;
    }
  };
  auto store_tile = [&]() {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      int ml = a_ml + (A_MFAST ? 0 : p * A_MSTEP);
      int kl = a_kl + (A_MFAST ? p * A_KSTEP : 0);
      As[kl][ml] = ra[p];
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      int nl = b_nl + (B_NFAST ? 0 : p * B_NSTEP);
      int kl = b_kl + (B_NFAST ? p * B_KSTEP : 0);
      Bs[kl][nl] = rb[p];
    }
  };

  d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const int li = lane & 15;  // row (A) / column (B) inside a 16x16 block
  const int lk = lane >> 4;  // k inside the 16x16x4 step
  const int wrow = wm * (BM / 2);
  const int wcol = wn * (BN / 2);

  if (kbeg < kend) load_tile(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    store_tile();
    __syncthreads();
    if (k0 + BK < kend) load_tile(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[ks * 4 + lk][wrow + i * 16 + li];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[ks * 4 + lk][wcol + j * 16 + li];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j],
                                                           acc[i][j], 0, 0, 0);
    }
  }

  // ---- epilogue -------------------------------------------------------------
  // v_mfma_f64_16x16x4 output layout (measured on gfx950): lane l, register
  // r -> row (l/16) + 4*r, column l%16.
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wcol + j * 16 + li;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wrow + i * 16 + lk + 4 * r;
        if (m >= M) continue;
        const double v = acc[i][j][r];
        if (ka.partial) {
          ka.partial[((int64_t)blockIdx.z * M + m) * N + n] = v;
        } else {
          double *dst = g.C + g.c_m[m] + g.c_n[n];
          *dst = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * (*dst);
        }
      }
    }
  }
}

__global__ void splitk_reduce_kernel(const double *P, int splits, int M, int N,
                                     double alpha, double beta, double *C,
                                     const int *c_m, const int *c_n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t MN = (int64_t)M * N;
  if (idx >= MN) return;
  double acc = 0.0;
  for (int z = 0; z < splits; ++z) acc += P[z * MN + idx];
  const int m = (int)(idx / N), n = (int)(idx % N);
  double *dst = C + c_m[m] + c_n[n];
  *dst = (beta == 0.0) ? alpha * acc : alpha * acc + beta * (*dst);
}

struct Plan {
  int BM, BN, splits, kps;
};

Plan choose(int M, int N, int K) {
  Plan p;
  if (K <= 0) {
    p.BM = p.BN = 64;
    p.splits = 1;
    p.kps = 0;
    return p;
  }
  const int64_t tiles128 = (int64_t)ceil_div(M, 128) * ceil_div(N, 128);
  if (M >= 128 && N >= 128 && tiles128 >= 200) {
    p.BM = p.BN = 128;
  } else {
    p.BM = p.BN = 64;
  }
  const int64_t tiles = (int64_t)ceil_div(M, p.BM) * ceil_div(N, p.BN);
  // Split K when the tile grid cannot fill the 256 CUs and K is long.
  p.splits = 1;
  if (tiles < 128 && K >= 4 * BK * 2) {
    int want = (int)((256 + tiles - 1) / tiles);
    int maxs = K / (4 * BK);
    p.splits = want < maxs ? want : maxs;
    if (p.splits < 1) p.splits = 1;
  }
  p.kps = ceil_div(K, p.splits);
  p.kps = ceil_div(p.kps, BK) * BK;
  p.splits = ceil_div(K, p.kps);
  return p;
}

template <int BM, int BN>
const char *launch_cfg(const KArgs &ka, dim3 grid, hipStream_t s, bool aff) {
  const bool am = ka.g.a_mfast, bn = ka.g.b_nfast;
#define TN_LAUNCH(AM, BNF, AF)                                              \
  do {                                                                      \
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, AM, BNF, AF>), grid,        \
                       dim3(256), 0, s, ka);                                \
    return BM == 128 ? "gemm_f64_kernel<128,128," #AM "," #BNF "," #AF ">"   \
                     : "gemm_f64_kernel<64,64," #AM "," #BNF "," #AF ">";     \
  } while (0)
  if (aff) {
    if (am && bn) TN_LAUNCH(true, true, true);
    else if (am) TN_LAUNCH(true, false, true);
    else if (bn) TN_LAUNCH(false, true, true);
    else TN_LAUNCH(false, false, true);
  } else {
    if (am && bn) TN_LAUNCH(true, true, false);
    else if (am) TN_LAUNCH(true, false, false);
    else if (bn) TN_LAUNCH(false, true, false);
    else TN_LAUNCH(false, false, false);
  }
#undef TN_LAUNCH
}

}  // namespace

size_t gemm_workspace_bytes(int M, int N, int K) {
  Plan p = choose(M, N, K);
  if (p.splits <= 1) return 0;
  return (size_t)p.splits * M * N * sizeof(double);
}

const char *gemm_f64(const GemmArgs &g, void *workspace, size_t ws_bytes,
                     hipStream_t s) {
  TN_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0, "gemm: negative size");
  if (g.M == 0 || g.N == 0) return "gemm_f64_empty";
  // K == 0 gives C = beta*C through a launch with an empty K range.
  Plan p = choose(g.M, g.N, g.K);
  KArgs ka;
  ka.g = g;
  ka.partial = nullptr;
  ka.k_per_split = p.kps > 0 ? p.kps : BK;
  int splits = p.splits;
  if (splits > 1) {
    size_t need = (size_t)splits * g.M * g.N * sizeof(double);
    if (workspace == nullptr || ws_bytes < need) {
      splits = 1;
      ka.k_per_split = g.K > 0 ? g.K : BK;
    } else {
      ka.partial = static_cast<double *>(workspace);
    }
  }
  if (g.K == 0) ka.k_per_split = 0;
  const bool aff = g.a_m_stride || g.a_k_stride || g.b_k_stride ||
                   g.b_n_stride;
  dim3 grid(ceil_div(g.N, p.BN), ceil_div(g.M, p.BM), splits);
  const char *name = (p.BM == 128) ? launch_cfg<128, 128>(ka, grid, s, aff)
                                   : launch_cfg<64, 64>(ka, grid, s, aff);
  TN_HIP_CHECK(hipGetLastError());
  if (ka.partial) {
    const int64_t MN = (int64_t)g.M * g.N;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(ceil_div(MN, 256)),
                       dim3(256), 0, s, ka.partial, splits, g.M, g.N, g.alpha,
                       g.beta, g.C, g.c_m, g.c_n);
    TN_HIP_CHECK(hipGetLastError());
  }
  return name;
}

void splitk_reduce(const double *P, int splits, int M, int N, double alpha,
                   double beta, double *C, const int *c_m, const int *c_n,
                   hipStream_t s) {
  const int64_t MN = (int64_t)M * N;
  if (MN == 0) return;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(ceil_div(MN, 256)), dim3(256),
                     0, s, P, splits, M, N, alpha, beta, C, c_m, c_n);
  TN_HIP_CHECK(hipGetLastError());
}

}  // namespace kern
}  // namespace tn
