// Device tensor type and the mptensor-like operation subset the CTMRG path
// uses (tensordot, transpose, reshape, slice, extend, max_abs, trace,
// multiply_vector), restated on HBM-resident FP64 row-major tensors.
//
// Semantics follow how TeNeS uses mptensor (src/iTPS/core/ctm.cpp and
// contract_itps_ctm/*.cpp call sites):
//   * tensordot(A, B, axA, axB): result legs = free legs of A (in order), then
//     free legs of B (in order);
//   * transpose(T, axes): numpy order, new leg k = old leg axes[k];
//   * reshape: row-major merge/split of legs;
//   * tensors have value semantics.  Here every operation returns a fresh
//     tensor and storage is never mutated once shared, so sharing storage
//     between copies is equivalent to the reference's deep copies.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace tn {

using Shape = std::vector<int64_t>;
using Axes = std::vector<int>;

class Context;

// Size-bucketed caching allocator.  Outside a fork all work of a Context
// runs on its main HIP stream, so a freed block may be handed to the next
// allocation without a device synchronisation: the stream orders the reuse
// after the last use.  While the context is forked over several streams,
// freed blocks are held back (deferred) until the join has ordered every
// stream's work before the main stream's.
class Allocator {
 public:
  explicit Allocator(int device) : device_(device) {}
  ~Allocator();
  void *alloc(size_t bytes);
  void free(void *p, size_t bytes);
  void release_cached();
  void set_deferred(bool on);
  size_t bytes_in_use() const { return in_use_; }
  size_t bytes_cached() const { return cached_; }

  // Arenas for captured graphs.  While an arena is active every allocation
  // comes from it (its own free list, else a new device block) and a block
  // of an arena always returns to that arena when freed, so the memory a
  // captured graph refers to never serves other work while the graph
  // exists.  Destroying an arena hands its blocks to the general pool (free
  // ones at once, the others when they are freed).
  int arena_create();
  void arena_activate(int id);  // -1: none
  void arena_destroy(int id);

 private:
  struct Arena {
    std::multimap<size_t, void *> free;
    std::vector<std::pair<size_t, void *>> deferred;
  };
  static size_t round(size_t bytes);
  static void *take(std::multimap<size_t, void *> &pool, size_t b);
  int device_;
  std::multimap<size_t, void *> free_;
  std::vector<std::pair<size_t, void *>> deferred_;
  bool defer_ = false;
  size_t in_use_ = 0, cached_ = 0;
  std::map<int, Arena> arenas_;
  std::map<void *, std::pair<int, size_t>> owner_;  // block -> (arena, bytes)
  int active_ = -1, next_arena_ = 0;
  std::mutex mu_;
};

struct Storage {
  double *ptr = nullptr;
  size_t bytes = 0;
  Context *ctx = nullptr;
  bool owned = true;
  ~Storage();
};

class Tensor {
 public:
  Tensor() = default;
  Shape shape;
  std::shared_ptr<Storage> st;

  int rank() const { return (int)shape.size(); }
  int64_t size() const {
    int64_t n = 1;
    for (auto d : shape) n *= d;
    return n;
  }
  double *data() const { return st ? st->ptr : nullptr; }
  Context *ctx() const { return st ? st->ctx : nullptr; }
  bool defined() const { return (bool)st; }
};

// Offset tables of one tensordot (device resident, cached per signature).
struct ContractPlan {
  int M = 0, N = 0, K = 0;
  Shape out_shape;
  int *d_tab = nullptr;  // a_m | a_k | b_k | b_n | c_m | c_n
  int a_m_off, a_k_off, b_k_off, b_n_off, c_m_off, c_n_off;
  bool a_mfast = false, b_nfast = false;
  int a_m_stride = 0, a_k_stride = 0, b_k_stride = 0, b_n_stride = 0;
  // GEMM launch shape picked by the autotuner (bm = 0: not tuned yet)
  mutable int tuned_bm = 0, tuned_splits = 0;
};

class Context {
 public:
  explicit Context(int device, hipStream_t stream = nullptr);
  ~Context();
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;

  int device() const { return device_; }
  // The stream all operations issue to: the main stream, or while forked the
  // side stream selected with use_lane().
  hipStream_t stream() const { return lanes_[cur_].s; }
  Allocator &allocator() { return alloc_; }

  // Fork/join over side streams.  fork(n): n side streams ("lanes") start
  // after all work issued so far on the main stream; use_lane(i) selects
  // lane i (i < 0: the main stream) for the operations that follow; join():
  // the main stream continues after all lanes' work and becomes current.
  // Each lane has its own scratch areas.  Calls that return host values
  // synchronise only the current stream.
  void fork(int n);
  void use_lane(int i);
  void join();
  // join() for error paths: never throws; the main stream waits for every
  // lane where the runtime still accepts the calls, and the context is left
  // on its main stream with frees no longer deferred in any case.
  void join_noexcept() noexcept;
  bool forked() const { return forked_; }
  int current_lane() const { return cur_ - 1; }  // -1: the main stream
  // Upper bound on the lanes the CTM moves fork into (1: no overlap, every
  // kernel on the main stream in issue order).
  int max_lanes() const { return max_lanes_; }
  void set_max_lanes(int n) { max_lanes_ = n < 1 ? 1 : (n > 16 ? 16 : n); }
  // Host run-ahead bound of the CTM driver: after issuing a move (or a pair
  // of moves) the host waits until at most n issued moves are unfinished on
  // the device (0: no bound).
  int max_moves_in_flight() const { return max_moves_; }
  void set_max_moves_in_flight(int n) { max_moves_ = n < 0 ? 0 : n; }
  void move_issued();

  // Scratch of the current stream (grows on demand, stream ordered).
  void *workspace(size_t bytes);
  double *scalars(int n);  // small device scalar area
  // Pinned host buffer for asynchronous device-to-host copies of small
  // results (slot < kPinnedSlots, at most 512 doubles each).  Slots below
  // kEagerSlots serve uncaptured work in issue order (eager_slot is the next
  // free one, reset when the pending statistics are collected); captured
  // sweeps own windows of slots above it (slot_window), so a replay never
  // writes a slot that other work in flight reads.  The last two slots
  // belong to the synchronous SVD and projector entry points.
  static constexpr int kPinnedSlots = 4096;
  static constexpr int kEagerSlots = 256;
  static constexpr int kSlotProjectorSync = kPinnedSlots - 2;
  static constexpr int kSlotSvdSync = kPinnedSlots - 1;
  int eager_slot = 0;
  // First slot of n free consecutive slots of the window range (-1 if none).
  int slot_window(int n);
  void release_slot_window(int first, int n);
  double *pinned(int slot);
  // The same slot's address for device code (kernels write it directly).
  double *pinned_device(int slot);
  // Projector SVDs whose singular values were left in pinned slots without
  // waiting for them (projector cut on the device): after the next
  // synchronisation of the stream collect_svd_stats() reads them, sets the
  // Ritz-value range of the warm-start entry (when `warm`) and adds the
  // Jacobi rounds to the statistics.
  struct PendingSvdStat {
    std::array<int64_t, 5> wkey;
    bool warm;
    int slot;
    int nall;
  };
  std::vector<PendingSvdStat> pending_svd;
  void collect_svd_stats();

  const ContractPlan &plan(const Shape &sa, const Shape &sb, const Axes &axa,
                           const Axes &axb, const Axes &perm);
  void sync();

  // Stream capture of the main stream (relaxed mode; side streams join
  // through the fork events).  While capturing nothing synchronises: a new
  // contraction plan (a synchronous upload) is an error, GEMM autotuning and
  // the event recycling of join() are skipped.  Every lane (the main stream
  // and all max_lanes() side streams) gets its own scalar area and an empty
  // workspace for the capture, allocated from the active allocator arena,
  // so the graph shares neither with uncaptured work nor with other graphs
  // (a workspace that later grows cannot pull memory out from under a
  // graph).  end_capture() returns the graph (nullptr if the capture was
  // invalidated), restores the lanes' own areas and hands the capture's
  // areas (pointer, bytes) to `blocks`: they belong to the graph.
  void begin_capture();
  hipGraph_t end_capture(std::vector<std::pair<void *, size_t>> &blocks);
  bool capturing() const { return capturing_; }
  // Captured CTM sweeps, one per environment (key: environment identity).
  struct SweepGraph {
    hipGraphExec_t exec = nullptr;
    int arena = -1;
    std::vector<std::pair<void *, size_t>> blocks;  // lane areas of the graph
    int slot_first = -1, slot_count = 0;             // its pinned slot window
    std::vector<uintptr_t> signature;
    std::vector<PendingSvdStat> pending;
    double gemm_flops = 0.0;
    int64_t gemm_calls = 0, svd_calls = 0;
  };
  std::map<int64_t, SweepGraph> sweep_graphs;
  int64_t graph_captures = 0, graph_replays = 0;
  // Waits for every lane, then releases the environment's graph.
  void drop_sweep_graph(int64_t environment);
  // Releases a graph's executable, lane areas, arena and slot window (the
  // caller has made sure no replay of it runs).
  void release_sweep_graph(SweepGraph &g);
  uint64_t next_rng_counter(uint64_t n) {
    uint64_t c = rng_counter_;
    rng_counter_ += n;
    return c;
  }
  uint64_t seed = 0x5EEDC7A1ull;

  std::string last_error;

  // accumulated statistics (for benchmarks)
  double gemm_flops = 0.0;
  int64_t gemm_calls = 0;
  // truncated SVDs through the Jacobi kernel and their round-robin rounds
  // (one round = one pass over p/2 disjoint column pairs)
  int64_t svd_calls = 0;
  int64_t jacobi_rounds = 0;
  // Range-finder start blocks of the CTM projectors (Ritz basis, its
  // singular values on the device, their max/min ratio once known), keyed by
  // (environment, move, coordinate, row, n).
  struct SvdStart {
    Tensor V, s;
    double range = 0.0;
  };
  std::map<std::array<int64_t, 5>, SvdStart> svd_start;
  void forget_svd_starts(int64_t environment);

  // Optional per-launch timing of the GEMM kernels with HIP events recorded
  // on this context's stream: per kernel instantiation, launch count, summed
  // device time and summed 2*M*N*K flops.
  struct KernelStat {
    int64_t launches = 0;
    double ms = 0.0;
    double flops = 0.0;
  };
  void set_profiling(bool on);
  bool profiling() const { return profiling_; }
  // Arithmetic of the contractions issued from now on: 0 = FP64 MFMA, 1 =
  // FP32 MFMA (operands rounded to FP32, FP32 accumulation), 2 = bf16 MFMA
  // with split operands (three hi/lo products, FP32 accumulation); FP64 data
  // in HBM in every mode.
  void set_contraction_mode(int mode) { mode_ = mode < 0 ? 0 : (mode > 2 ? 2 : mode); }
  int contraction_mode() const { return mode_; }
  void set_fp32_contractions(bool on) { mode_ = on ? 1 : 0; }
  bool fp32_contractions() const { return mode_ == 1; }
  // The HBM-bound pairs of the enlarged corners and Calc_Next_eT as one
  // double-layer site kernel (FP64, double-layer networks; on by default);
  // off: the two gather-GEMMs of the reference's contraction order.
  void set_fused_site(bool on) { fused_site_ = on; }
  bool fused_site() const { return fused_site_; }
  // Single-layer networks: while on, bra() of a site tensor is the unit
  // tensor of shape (1,1,1,1,1), so the double-layer contractions of the
  // CTM moves and of the environment contractions become those of the
  // single-layer network of the ket tensors alone (TPOs traced over their
  // physical legs; edges (chi, chi, D, 1)).
  void set_single_layer(bool on) { single_layer_ = on; }
  bool single_layer() const { return single_layer_; }
  // GEMM autotuning: while on, the first contraction of every plan (shapes
  // and legs) with a GEMM of at least kAutotuneFlops times each candidate
  // launch shape (tile size, K splits) on the stream and keeps the fastest
  // for that plan; later calls use it.
  void set_autotune(bool on) { autotune_ = on; }
  bool autotune() const { return autotune_; }
  static constexpr double kAutotuneFlops = 1e8;
  // Stream-K launches for the FP64 GEMMs that qualify (kern::gemm_stream_k):
  // 128 x 128 tiles whose grid leaves a partial round.  Off by default
  // (slower than the data-parallel grid at 2500^3 on this part); the
  // autotuner also tries it.
  void set_stream_k(bool on) { stream_k_ = on; }
  bool stream_k() const { return stream_k_; }
  void profile_begin(hipEvent_t *a);
  void profile_end(hipEvent_t a, const char *name, double flops);
  std::map<std::string, KernelStat> kernel_stats();  // flushes pending
  void reset_kernel_stats();

 private:
  struct Lane {
    hipStream_t s = nullptr;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    double *scal = nullptr;
  };
  int device_;
  bool own_stream_ = false;
  Allocator alloc_;
  std::vector<Lane> lanes_;  // [0] = main stream, then the side streams
  void add_lanes(int n);     // creates side streams until there are n
  int cur_ = 0;
  int nforked_ = 0;
  int max_lanes_ = 8;
  int max_moves_ = 0;
  bool capturing_ = false;
  struct LaneAreas {
    void *ws;
    size_t ws_bytes;
    double *scal;
  };
  std::vector<LaneAreas> saved_lanes_;  // the lanes' own areas during a capture
  std::vector<bool> slot_used_;         // window range, kEagerSlots upward
  std::vector<hipEvent_t> inflight_;  // end of each issued move, oldest first
  bool forked_ = false;
  hipEvent_t fork_event_ = nullptr;
  double *pinned_ = nullptr;
  double *pinned_dev_ = nullptr;
  std::map<std::string, ContractPlan> plans_;
  uint64_t rng_counter_ = 0;
  int mode_ = 0;
  bool fused_site_ = true;
  bool single_layer_ = false;
  bool autotune_ = false;
  bool stream_k_ = false;
  bool profiling_ = false;
  struct Pending {
    hipEvent_t a, b;
    std::string name;
    double flops;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> event_pool_;
  std::vector<hipEvent_t> joined_events_;
  std::map<std::string, KernelStat> kstats_;
  hipEvent_t get_event();
  void flush_pending();
};

// A fork of a context that cannot outlive its scope: end() joins on the
// normal path; leaving the scope by an exception joins with
// join_noexcept(), so an error inside the fork never leaves later work of
// the context on a side stream or its frees deferred.
class ForkScope {
 public:
  ForkScope(Context &ctx, int lanes, bool on) : ctx_(ctx), active_(on) {
    if (active_) ctx_.fork(lanes);
  }
  void end() {
    if (active_) {
      active_ = false;
      ctx_.join();
    }
  }
  ~ForkScope() {
    if (active_) ctx_.join_noexcept();
  }
  ForkScope(const ForkScope &) = delete;
  ForkScope &operator=(const ForkScope &) = delete;

 private:
  Context &ctx_;
  bool active_;
};

// Sets the contraction precision of a context for a scope.
class ContractionPrecision {
 public:
  ContractionPrecision(Context &ctx, int mode) : ctx_(ctx), old_(ctx.contraction_mode()) {
    ctx.set_contraction_mode(mode);
  }
  ~ContractionPrecision() { ctx_.set_contraction_mode(old_); }
  ContractionPrecision(const ContractionPrecision &) = delete;
  ContractionPrecision &operator=(const ContractionPrecision &) = delete;

 private:
  Context &ctx_;
  int old_;
};

// Sets the single-layer mode of a context for a scope.
class SingleLayer {
 public:
  SingleLayer(Context &ctx, bool on) : ctx_(ctx), old_(ctx.single_layer()) {
    ctx.set_single_layer(on);
  }
  ~SingleLayer() { ctx_.set_single_layer(old_); }
  SingleLayer(const SingleLayer &) = delete;
  SingleLayer &operator=(const SingleLayer &) = delete;

 private:
  Context &ctx_;
  bool old_;
};

// ---- construction / transfer ----------------------------------------------
Tensor empty(Context &ctx, const Shape &shape);
Tensor zeros(Context &ctx, const Shape &shape);
Tensor from_host(Context &ctx, const Shape &shape, const double *host);
void to_host(const Tensor &t, double *host);  // synchronous
// Wrap caller-owned device memory (not freed by the library).
Tensor borrow(Context &ctx, const Shape &shape, double *dev);
Tensor clone(const Tensor &t);

// ---- mptensor-subset -------------------------------------------------------
Tensor reshape(const Tensor &t, const Shape &shape);  // shares storage
Tensor transpose(const Tensor &t, const Axes &axes);
Tensor tensordot(const Tensor &A, const Tensor &B, const Axes &axA,
                 const Axes &axB);
// tensordot followed by a transpose of the result, fused into the store.
Tensor tensordot_t(const Tensor &A, const Tensor &B, const Axes &axA,
                   const Axes &axB, const Axes &perm);
// C(+)= tensordot into an existing tensor with the given output perm.
void tensordot_into(const Tensor &A, const Tensor &B, const Axes &axA,
                    const Axes &axB, const Axes &perm, Tensor &C, double alpha,
                    double beta);
Tensor slice(const Tensor &t, int axis, int64_t begin, int64_t end);
Tensor extend(const Tensor &t, const Shape &shape);  // zero padding
Tensor conj(const Tensor &t);                         // real: identity
// The bra copy of a site tensor (l, u, r, d, p) in a double-layer network:
// conj(t), or the unit (1,1,1,1,1) tensor in single-layer mode (t must then
// have a physical dimension of one).
Tensor bra(const Tensor &t);
// Full contraction of A and B over all legs (A leg axA[k] with B leg axB[k])
// to a scalar (mptensor trace(A, B, axA, axB)).  Synchronous.
double trace(const Tensor &A, const Tensor &B, const Axes &axA,
             const Axes &axB);
// Same, result left on the device in *out.
void trace_dev(const Tensor &A, const Tensor &B, const Axes &axA,
               const Axes &axB, double *out);
double max_abs(const Tensor &t);  // synchronous
// t /= max_abs(t), computed and applied on the device.
void normalize_max_abs(Tensor &t);
// Scale leg `axis` by the device vector v (length shape[axis]).
void multiply_vector_dev(Tensor &t, const double *v_dev, int axis);
void scale_inplace(Tensor &t, double a);
Tensor add(const Tensor &a, const Tensor &b, double alpha = 1.0,
           double beta = 1.0);

std::string shape_str(const Shape &s);

}  // namespace tn
