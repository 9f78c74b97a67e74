// FP64 gather-GEMM for gfx950 (CDNA4) on v_mfma_f64_16x16x4_f64, with an
// FP32-compute variant on v_mfma_f32_16x16x4_f32 (same tile shape, twice the
// matrix rate: operands rounded to FP32 when staged into LDS, FP32
// accumulation, FP64 in HBM on both sides).
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
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int BK = 16;

// Compute-type traits: accumulator vector and the 16x16x4 MFMA.
template <class CT>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ d4 run(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ f4 run(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

struct KArgs {
  GemmArgs g;
  int k_per_split;
  double *partial;  // split-K partials [z][M][N] or nullptr
  // Stream-K (gemm_kernel<..., SK = true>, 128 x 128 FP64 tiles, a grid of
  // kSkRanks workgroups): the first sk_dp tiles of the logical tile order run
  // whole, kSkRanks at a time; the K iterations (BK each) of the remaining
  // sk_tiles tiles are split evenly over the ranks.  A tile split between
  // ranks is summed by the last of them to finish, from sk_S partial slots
  // per rank (sk_part, 64 x 256 doubles each) and one arrival counter per
  // tile (sk_cnt, zero on entry, zeroed again by that last rank).
  int sk_dp = 0, sk_tiles = 0, sk_nk = 0, sk_S = 0;
  int64_t sk_iters = 0;
  double *sk_part = nullptr;
  int *sk_cnt = nullptr;
};

// XCD-aware tile order.  Workgroups are dealt to the 8 XCDs round-robin by
// linear id, and each XCD has its own L2; so the ids of XCD x (x, x+8, ...)
// are given a contiguous range of the logical tile order, and that order
// walks the output in groups of kGroupM tile rows (column by column inside
// a group).  The tiles running at the same time on one XCD then share a few
// A row panels and B column panels instead of touching all of them.
constexpr int kXcds = 8;
constexpr int kGroupM = 8;
__device__ __forceinline__ void tile_coords(int t, int nbm, int nbn, int &bm, int &bn);
__device__ __forceinline__ void tile_of(int id, int nbm, int nbn, int &bm, int &bn) {
  const int ntiles = nbm * nbn;
  const int xcd = id % kXcds, slot = id / kXcds;
  const int base = ntiles / kXcds, extra = ntiles % kXcds;
  const int t = xcd * base + min(xcd, extra) + slot;  // bijection on [0, ntiles)
  tile_coords(t, nbm, nbn, bm, bn);
}
// Logical tile t -> (tile row, tile column): groups of kGroupM tile rows,
// column by column inside a group.
__device__ __forceinline__ void tile_coords(int t, int nbm, int nbn, int &bm, int &bn) {
  const int per_group = kGroupM * nbn;
  const int g = t / per_group;
  const int first_m = g * kGroupM;
  const int gsize = min(nbm - first_m, kGroupM);
  const int r = t - g * per_group;
  bm = first_m + r % gsize;
  bn = r / gsize;
}

constexpr int kSkRanks = 512;  // stream-K grid: two workgroups per CU

template <class CT, int BM, int BN, bool A_MFAST, bool B_NFAST, bool KA, bool KB,
          bool SK = false>
__global__ __launch_bounds__(256, SK ? 2 : 1) void gemm_kernel(KArgs ka) {
  // KA / KB: the contracted-index offsets of A / B form an arithmetic
  // progression (computed), otherwise they come from the table, prefetched
  // one K tile ahead so that the tile loads never wait on a table read.
  const GemmArgs &g = ka.g;
  constexpr int TM = BM / 32;
  constexpr int TN = BN / 32;
  constexpr int A_PER = BM * BK / 256;
  constexpr int B_PER = BN * BK / 256;
  constexpr int A_NK = A_MFAST ? A_PER : 1;  // distinct k per thread and tile
  constexpr int B_NK = B_NFAST ? B_PER : 1;
  __shared__ CT As[2][BK][BM + 1];  // double-buffered K tiles
  __shared__ CT Bs[2][BK][BN + 1];
  typedef typename Mfma<CT>::acc_t acc_t;

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int M = g.M, N = g.N;

  // the work item: output tile (m0, n0) over the K range [kbeg, kend)
  int kbeg, kend, m0, n0;

  // ---- loader geometry ----------------------------------------------------
  // Element p of this thread in the A tile (BM x BK) is e = t + 256 p of the
  // tile in the lane-fast order: MFAST (consecutive lanes walk m, coalesced
  // when A's m stride is 1) e = ml + BM kl, else (lanes walk k) e = kl + BK ml.
  // Same for B with n.  (ml, kl) and the outer offsets are fixed for the
  // whole K loop and kept in registers.
  int a_ml[A_PER], a_kl[A_PER], a_moff[A_PER];
  bool a_mok[A_PER];
  int b_nl[B_PER], b_kl[B_PER], b_noff[B_PER];
  bool b_nok[B_PER];
  auto setup = [&]() {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      const int e = t + 256 * p;
      a_ml[p] = A_MFAST ? e % BM : e / BK;
      a_kl[p] = A_MFAST ? e / BM : e % BK;
      const int m = m0 + a_ml[p];
      a_mok[p] = m < M;
      a_moff[p] = a_mok[p] ? (g.a_m_stride ? m * g.a_m_stride : g.a_m[m]) : 0;
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      const int e = t + 256 * p;
      b_nl[p] = B_NFAST ? e % BN : e / BK;
      b_kl[p] = B_NFAST ? e / BN : e % BK;
      const int n = n0 + b_nl[p];
      b_nok[p] = n < N;
      b_noff[p] = b_nok[p] ? (g.b_n_stride ? n * g.b_n_stride : g.b_n[n]) : 0;
    }
  };

  // k of load p of this thread within the tile starting at k0 (for the
  // k-fast layouts all loads of a thread share one k)
  auto a_k_of = [&](int k0, int p) { return k0 + a_kl[A_MFAST ? p : 0]; };
  auto b_k_of = [&](int k0, int p) { return k0 + b_kl[B_NFAST ? p : 0]; };
  int ka_off[A_NK], kb_off[B_NK];  // table offsets of the next tile (!KA / !KB)
  auto fetch_koff = [&](int k0) {
    if (!KA) {
#pragma unroll
      for (int p = 0; p < A_NK; ++p) {
        const int kk = a_k_of(k0, p);
        ka_off[p] = g.a_k[kk < kend ? kk : kbeg];
      }
    }
    if (!KB) {
#pragma unroll
      for (int p = 0; p < B_NK; ++p) {
        const int kk = b_k_of(k0, p);
        kb_off[p] = g.b_k[kk < kend ? kk : kbeg];
      }
    }
  };

  // Branch-free tile loads: out-of-range elements read offset 0 (a valid
  // address) and are replaced by zero only when the tile is stored to LDS,
  // so all loads of a tile stay in flight during the MFMA work.  Loads and
  // stores are issued in four parts (part q: elements p with p % 4 == q) so
  // that the main loop can spread them between its MFMA steps.
  double ra[A_PER], rb[B_PER];
  bool oka[A_PER], okb[B_PER];
  auto load_part = [&](int k0, int q) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      if (A_PER >= 4 && p % 4 != q) continue;
      if (A_PER < 4 && q != 0) continue;
      const int kk = a_k_of(k0, p);
      oka[p] = a_mok[p] && kk < kend;
      const int koff = KA ? kk * g.a_k_stride : ka_off[A_MFAST ? p : 0];
      const int off = a_moff[p] + koff;
      ra[p] = g.A[oka[p] ? off : 0];
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      if (B_PER >= 4 && p % 4 != q) continue;
      if (B_PER < 4 && q != 0) continue;
      const int kk = b_k_of(k0, p);
      okb[p] = b_nok[p] && kk < kend;
      const int koff = KB ? kk * g.b_k_stride : kb_off[B_NFAST ? p : 0];
      const int off = b_noff[p] + koff;
      rb[p] = g.B[okb[p] ? off : 0];
    }
  };
  auto store_part = [&](int buf, int q) {
#pragma unroll
    for (int p = 0; p < A_PER; ++p) {
      if (A_PER >= 4 && p % 4 != q) continue;
      if (A_PER < 4 && q != 0) continue;
      As[buf][a_kl[p]][a_ml[p]] = (CT)(oka[p] ? ra[p] : 0.0);
    }
#pragma unroll
    for (int p = 0; p < B_PER; ++p) {
      if (B_PER >= 4 && p % 4 != q) continue;
      if (B_PER < 4 && q != 0) continue;
      Bs[buf][b_kl[p]][b_nl[p]] = (CT)(okb[p] ? rb[p] : 0.0);
    }
  };

  acc_t acc[TM][TN];
  const int li = lane & 15;  // row (A) / column (B) inside a 16x16 block
  const int lk = lane >> 4;  // k inside the 16x16x4 step
  const int wrow = wm * (BM / 2);
  const int wcol = wn * (BN / 2);

  // acc = the (m0, n0) tile's product over [kbeg, kend); ends with a barrier
  // after the last LDS read, so the next work item may refill the buffers
  auto mainloop = [&]() {
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  if (BM == 128) {
    // 128 x 128 tiles: tile t+1 is loaded into registers while tile t is
    // multiplied from LDS buffer t&1, then stored into the other buffer; one
    // barrier per K tile.  (The interleaved pipeline below needs ~45 more
    // registers here, which costs the second workgroup per CU.)
    if (kbeg < kend) {
      fetch_koff(kbeg);
#pragma unroll
      for (int q = 0; q < 4; ++q) load_part(kbeg, q);
#pragma unroll
      for (int q = 0; q < 4; ++q) store_part(0, q);
      fetch_koff(kbeg + BK);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool next = k0 + BK < kend;
      if (next) {
#pragma unroll
        for (int q = 0; q < 4; ++q) load_part(k0 + BK, q);
        fetch_koff(k0 + 2 * BK);
      }
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        CT a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = As[buf][ks * 4 + lk][wrow + i * 16 + li];
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = Bs[buf][ks * 4 + lk][wcol + j * 16 + li];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<CT>::run(a[i], b[j], acc[i][j]);
      }
      if (next) {
#pragma unroll
        for (int q = 0; q < 4; ++q) store_part(buf ^ 1, q);
      }
      __syncthreads();
      buf ^= 1;
    }
  } else {
  // Pipeline: tile t is multiplied from LDS buffer t&1 while tile t+1 (in
  // registers since the previous tile) is stored into the other buffer and
  // tile t+2 is loaded into the freed registers, a quarter of each after
  // each of the four MFMA steps, so the MFMA pipe keeps running while the
  // memory instructions issue; one barrier per K tile.  The loop body has
  // no branches: past the end of the K range the loads are masked (zeros
  // stored into the buffer no later tile reads).
  if (kbeg < kend) {
    fetch_koff(kbeg);
#pragma unroll
    for (int q = 0; q < 4; ++q) load_part(kbeg, q);
#pragma unroll
    for (int q = 0; q < 4; ++q) store_part(0, q);
    fetch_koff(kbeg + BK);
#pragma unroll
    for (int q = 0; q < 4; ++q) load_part(kbeg + BK, q);
    fetch_koff(kbeg + 2 * BK);
  }
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      CT a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[buf][ks * 4 + lk][wrow + i * 16 + li];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[buf][ks * 4 + lk][wcol + j * 16 + li];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<CT>::run(a[i], b[j], acc[i][j]);
      store_part(buf ^ 1, ks);
      load_part(k0 + 2 * BK, ks);
      // keep the parts in program order: the compiler would otherwise hoist
      // all stores (and their waits on the loads) to the top of the tile
      __builtin_amdgcn_sched_barrier(0);
    }
    fetch_koff(k0 + 3 * BK);
    __syncthreads();
    buf ^= 1;
  }
  }
  };

  // ---- epilogue -------------------------------------------------------------
  // Output layouts (lane l, register r, column l%16): v_mfma_f64_16x16x4
  // row (l/16) + 4*r (measured on gfx950); v_mfma_f32_16x16x4 row
  // 4*(l/16) + r.
  auto epilogue = [&](int z) {
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wcol + j * 16 + li;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = sizeof(CT) == 8 ? lk + 4 * r : 4 * lk + r;
        const int m = m0 + wrow + i * 16 + row;
        if (m >= M) continue;
        const double v = (double)acc[i][j][r];
        if (ka.partial) {
          ka.partial[((int64_t)z * M + m) * N + n] = v;
        } else {
          double *dst = g.C + g.c_m[m] + g.c_n[n];
          *dst = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * (*dst);
        }
      }
    }
  }
  };

  if (!SK) {
    int bm, bn;
    tile_of(blockIdx.y * gridDim.x + blockIdx.x, gridDim.y, gridDim.x, bm, bn);
    kbeg = blockIdx.z * ka.k_per_split;
    kend = min(g.K, kbeg + ka.k_per_split);
    m0 = bm * BM;
    n0 = bn * BN;
    setup();
    mainloop();
    epilogue(blockIdx.z);
    return;
  }

  // ---- stream-K -------------------------------------------------------------
  // Rank r of this workgroup: the ranks of one XCD (ids x, x+8, ...) are
  // consecutive, so its data-parallel tiles and its stretch of the split
  // iterations are neighbours in the logical tile order (shared panels in
  // that XCD's L2).  One loop over the rank's work items (its data-parallel
  // tiles, then its pieces of split tiles) so the tile code is inlined once.
  __shared__ int s_last;
  const int nbm = (M + BM - 1) / BM, nbn = (N + BN - 1) / BN;
  const int nk = ka.sk_nk;
  const int rank = (blockIdx.x % kXcds) * (kSkRanks / kXcds) + blockIdx.x / kXcds;
  const int64_t I = ka.sk_iters;
  const int64_t beg = (int64_t)rank * I / kSkRanks;
  const int64_t end = (int64_t)(rank + 1) * I / kSkRanks;
  // the rank whose range holds split iteration x (largest r, r I / R <= x)
  auto rank_of = [&](int64_t x) { return (int)(((x + 1) * kSkRanks + I - 1) / I - 1); };
  constexpr int NREG = TM * TN * 4;
  int tt = rank;
  int64_t x = beg;
  while (tt < ka.sk_dp || x < end) {
    const bool dp = tt < ka.sk_dp;
    int tile, ts = 0, kb = 0, ke = nk;
    if (dp) {
      tile = tt;
      tt += kSkRanks;
    } else {
      ts = (int)(x / nk);
      kb = (int)(x - (int64_t)ts * nk);
      ke = (int)min((int64_t)nk, end - (int64_t)ts * nk);
      tile = ka.sk_dp + ts;
      x = (int64_t)ts * nk + ke;
    }
    int bm, bn;
    tile_coords(tile, nbm, nbn, bm, bn);
    m0 = bm * BM;
    n0 = bn * BN;
    kbeg = kb * BK;
    kend = min(g.K, ke * BK);
    setup();
    mainloop();
    const bool shared = !dp && !(kb == 0 && ke == nk);
    if (shared) {
      // this rank's share of tile ts: its partial slot, then the arrival
      // count; the last of the tile's ranks to arrive writes the tile
      const int first = rank_of((int64_t)ts * nk);
      const int last = rank_of((int64_t)ts * nk + nk - 1);
      auto slot_ptr = [&](int r) {
        const int64_t rb = (int64_t)r * I / kSkRanks;
        const int64_t sl = (int64_t)r * ka.sk_S + (ts - (int)(rb / nk));
        return ka.sk_part + sl * NREG * 256;
      };
      double *mine = slot_ptr(rank);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            mine[((i * TN + j) * 4 + r) * 256 + t] = (double)acc[i][j][r];
      __threadfence();
      __syncthreads();
      if (t == 0) s_last = atomicAdd(ka.sk_cnt + ts, 1) == last - first;
      __syncthreads();
      if (!s_last) continue;
      __threadfence();
      // sum the shares from their slots (this rank's own included) in rank
      // order: the same sum whichever rank arrives last
      for (int r = first; r <= last; ++r) {
        const double *src = slot_ptr(r);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const CT v = (CT)src[((i * TN + j) * 4 + q) * 256 + t];
              acc[i][j][q] = r == first ? v : acc[i][j][q] + v;
            }
      }
      if (t == 0) ka.sk_cnt[ts] = 0;  // ready for the next launch
    }
    epilogue(0);
  }
}

// Zero the stream-K arrival counters (a kernel, not a memset, so a captured
// graph replays it correctly).
__global__ void zero_ints_kernel(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
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
  bool sk = false;  // stream-K (128 x 128 tiles, kSkRanks workgroups)
};

// Stream-K layout of an M x N x K product on 128 x 128 tiles: whole rounds
// of kSkRanks tiles run data-parallel, the K iterations of the remaining
// tiles are split over the kSkRanks ranks.  Used when the remainder is at
// least half a round (below that a split tile has more than ~4 shares; the
// classic grid of two workgroups per CU leaves some CUs with one tile and
// others with two) and K is long enough (>= 64 iterations) for the shares to
// amortise their partial tiles.
struct SkLayout {
  int dp = 0, tiles = 0, nk = 0, S = 0;
  int64_t iters = 0;
  size_t part_bytes = 0, bytes = 0;
};
bool sk_layout(int M, int N, int K, SkLayout &L) {
  if (M < 128 || N < 128 || K <= 0) return false;
  const int64_t ntiles = (int64_t)ceil_div(M, 128) * ceil_div(N, 128);
  const int nk = ceil_div(K, BK);
  const int64_t dp = (ntiles / kSkRanks) * kSkRanks;
  const int64_t rest = ntiles - dp;
  if (rest < kSkRanks / 2 || nk < 64 || ntiles > (1 << 20)) return false;
  L.dp = (int)dp;
  L.tiles = (int)rest;
  L.nk = nk;
  L.iters = rest * nk;
  const int64_t lmax = ceil_div(L.iters, (int64_t)kSkRanks);
  L.S = (int)ceil_div(lmax, (int64_t)nk) + 1;
  L.part_bytes = (size_t)kSkRanks * L.S * 64 * 256 * sizeof(double);
  L.bytes = L.part_bytes + (size_t)L.tiles * sizeof(int);
  return true;
}

Plan choose(int M, int N, int K) {
  Plan p;
  if (K <= 0) {
    p.BM = p.BN = 64;
    p.splits = 1;
    p.kps = 0;
    return p;
  }
  // Large outputs with a long K: 128 x 128 tiles.  (160 x 160 tiles would
  // fill the 256 CUs in one round at 2500 x 2500, but measured 1.03 ms
  // against 0.78 ms: their 25 accumulator tiles per wave leave too few
  // registers.)  With K of at most 8 tiles the 64 x 64 grid, four times the
  // workgroups at twice the occupancy, hides more of the short main loop's
  // prologue and epilogue: 2500 x 2500 x 100 in 47 us against 53 us.
  const int64_t tiles128 = (int64_t)ceil_div(M, 128) * ceil_div(N, 128);
  p.BM = p.BN = (M >= 128 && N >= 128 && tiles128 >= 200 && K > 8 * BK) ? 128 : 64;
  const int64_t tiles = (int64_t)ceil_div(M, p.BM) * ceil_div(N, p.BN);
  // Split K when the tile grid cannot fill the 256 CUs and K is long: the
  // split count s that minimises the rounds of 256 workgroups per unit of
  // K work, ceil(tiles s / 256) / s, among those giving at least two
  // workgroups per CU (64 x 64 tiles run up to four per CU; one alone
  // leaves the MFMA pipe idle at every barrier), at least four K tiles per
  // split, fewest splits on a tie.  2500 x 128 x 2500 (80 tiles): 16 splits,
  // 51.5 us against 64.5 us at the 4 that merely reach 256 workgroups.
  p.splits = 1;
  if (tiles < 128 && K >= 4 * BK * 2) {
    const int maxs = std::max(1, K / (4 * BK));
    double best = 1e30;
    for (int sp = 1; sp <= std::min(maxs, 64); ++sp) {
      const int64_t wg = tiles * sp;
      if (wg < 512 && sp < maxs && sp < 64) continue;
      const double cost = (double)ceil_div(wg, (int64_t)256) / sp;
      if (cost < best - 1e-12) {
        best = cost;
        p.splits = sp;
      }
    }
  }
  p.kps = ceil_div(K, p.splits);
  p.kps = ceil_div(p.kps, BK) * BK;
  p.splits = ceil_div(K, p.kps);
  // stream-K only on request (GemmPlan{128, 0}): at 2500 x 2500 x 2500 it
  // measured 0.82 ms against 0.74 ms for the data-parallel grid
  return p;
}

// Static name of an instantiation: "gemm_f64_kernel<BM,BN,A_MFAST,B_NFAST,
// KA,KB>" or "gemm_f32_kernel<...>" (used by the per-launch timing report).
template <class CT, int BM, bool AM, bool BNF, bool KAA, bool KBA>
const char *kernel_name() {
  static const std::string n = std::string(sizeof(CT) == 8 ? "gemm_f64_kernel<" : "gemm_f32_kernel<") +
                               std::to_string(BM) + "," + std::to_string(BM) +
                               (AM ? ",true" : ",false") +
                               (BNF ? ",true" : ",false") +
                               (KAA ? ",true" : ",false") +
                               (KBA ? ",true>" : ",false>");
  return n.c_str();
}

template <class CT, int BM, int BN, bool AM, bool BNF>
const char *launch_k(const KArgs &ka, dim3 grid, hipStream_t s) {
  const bool ka_aff = ka.g.a_k_stride != 0, kb_aff = ka.g.b_k_stride != 0;
#define TN_LAUNCH(KAA, KBA)                                                  \
  do {                                                                       \
    hipLaunchKernelGGL((gemm_kernel<CT, BM, BN, AM, BNF, KAA, KBA>), grid,   \
                       dim3(256), 0, s, ka);                                 \
    return kernel_name<CT, BM, AM, BNF, KAA, KBA>();                         \
  } while (0)
  if (ka_aff && kb_aff) TN_LAUNCH(true, true);
  if (ka_aff) TN_LAUNCH(true, false);
  if (kb_aff) TN_LAUNCH(false, true);
  TN_LAUNCH(false, false);
#undef TN_LAUNCH
}

template <class CT, int BM, int BN>
const char *launch_cfg(const KArgs &ka, dim3 grid, hipStream_t s) {
  const bool am = ka.g.a_mfast, bn = ka.g.b_nfast;
  if (am && bn) return launch_k<CT, BM, BN, true, true>(ka, grid, s);
  if (am) return launch_k<CT, BM, BN, true, false>(ka, grid, s);
  if (bn) return launch_k<CT, BM, BN, false, true>(ka, grid, s);
  return launch_k<CT, BM, BN, false, false>(ka, grid, s);
}

// The heuristic plan, or the given launch shape with its K range rounded
// to whole tiles (at most one split per K tile).
Plan plan_for(int M, int N, int K, const GemmPlan *gp) {
  if (gp == nullptr || gp->bm == 0 || K <= 0) return choose(M, N, K);
  Plan p;
  p.BM = p.BN = gp->bm == 128 ? 128 : 64;
  SkLayout L;
  if (p.BM == 128 && gp->splits == 0 && sk_layout(M, N, K, L)) {
    p.splits = 1;
    p.kps = ceil_div(K, BK) * BK;
    p.sk = true;
    return p;
  }
  const int maxs = std::max(1, ceil_div(K, BK));
  p.splits = std::min(std::max(gp->splits, 1), maxs);
  p.kps = ceil_div(ceil_div(K, p.splits), BK) * BK;
  p.splits = ceil_div(K, p.kps);
  return p;
}

}  // namespace

bool gemm_stream_k(int M, int N, int K) {
  SkLayout L;
  return sk_layout(M, N, K, L);
}

std::vector<GemmPlan> gemm_candidates(int M, int N, int K) {
  std::vector<GemmPlan> out;
  const Plan h = choose(M, N, K);
  auto add = [&](int bm, int splits) {
    if (bm == 128 && (M < 128 || N < 128)) return;
    GemmPlan g{bm, splits};
    const Plan r = plan_for(M, N, K, &g);
    for (const auto &o : out) {
      const Plan e = plan_for(M, N, K, &o);
      if (e.BM == r.BM && e.splits == r.splits && e.sk == r.sk) return;
    }
    out.push_back(g);
  };
  add(h.BM, h.splits);
  add(64, 1);
  add(128, 1);
  if (gemm_stream_k(M, N, K)) add(128, 0);  // splits 0 with bm 128: stream-K
  if (K >= 8 * BK) {
    add(64, std::max(2, h.splits * 2));
    add(64, std::max(2, h.splits / 2));
    add(128, std::max(2, h.splits));
  }
  return out;
}

size_t gemm_workspace_bytes(int M, int N, int K, const GemmPlan *plan) {
  Plan p = plan_for(M, N, K, plan);
  SkLayout L;
  if (p.sk && sk_layout(M, N, K, L)) return L.bytes;
  if (p.splits <= 1) return 0;
  return (size_t)p.splits * M * N * sizeof(double);
}

const char *gemm_f64(const GemmArgs &g, void *workspace, size_t ws_bytes,
                     hipStream_t s, const GemmPlan *plan) {
  TN_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0, "gemm: negative size");
  if (g.M == 0 || g.N == 0) return "gemm_f64_empty";
  // K == 0 gives C = beta*C through a launch with an empty K range.
  Plan p = plan_for(g.M, g.N, g.K, plan);
  if (g.bf16x3)
    return gemm_bf16x3(g, p.BM, p.splits, p.kps, workspace, ws_bytes, s);
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
  // stream-K: FP64, the row-major R = LT * LB operand layout (the
  // instantiation built for it), the workspace for its partial tiles
  SkLayout L;
  if (p.sk && !g.fp32 && !g.a_mfast && g.b_nfast && g.a_k_stride != 0 && g.b_k_stride != 0 &&
      sk_layout(g.M, g.N, g.K, L) && workspace != nullptr && ws_bytes >= L.bytes) {
    ka.sk_dp = L.dp;
    ka.sk_tiles = L.tiles;
    ka.sk_nk = L.nk;
    ka.sk_S = L.S;
    ka.sk_iters = L.iters;
    ka.sk_part = static_cast<double *>(workspace);
    ka.sk_cnt = reinterpret_cast<int *>(static_cast<char *>(workspace) + L.part_bytes);
    ka.k_per_split = g.K;
    hipLaunchKernelGGL(zero_ints_kernel, dim3(ceil_div(L.tiles, 256)), dim3(256), 0, s,
                       ka.sk_cnt, L.tiles);
    hipLaunchKernelGGL((gemm_kernel<double, 128, 128, false, true, true, true, true>),
                       dim3(kSkRanks), dim3(256), 0, s, ka);
    TN_HIP_CHECK(hipGetLastError());
    static const std::string name = std::string(kernel_name<double, 128, false, true, true, true>()) + "+sk";
    return name.c_str();
  }
  dim3 grid(ceil_div(g.N, p.BN), ceil_div(g.M, p.BM), splits);
  const char *name =
      g.fp32 ? ((p.BM == 128) ? launch_cfg<float, 128, 128>(ka, grid, s)
                              : launch_cfg<float, 64, 64>(ka, grid, s))
             : ((p.BM == 128) ? launch_cfg<double, 128, 128>(ka, grid, s)
                              : launch_cfg<double, 64, 64>(ka, grid, s));
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
