// Device tensor operations: allocator, context, tensordot planning.
#include "tensor.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <sstream>

#include "kernels.hpp"

namespace tn {

// ---------------------------------------------------------------------------
// Allocator
// ---------------------------------------------------------------------------
size_t Allocator::round(size_t bytes) {
  if (bytes < 512) return 512;
  // four size classes per power of two
  size_t p = 1;
  while (p < bytes) p <<= 1;
  const size_t q = p >> 3;  // eighth of the next power of two
  return ((bytes + q - 1) / q) * q;
}

void *Allocator::take(std::multimap<size_t, void *> &pool, size_t b) {
  auto it = pool.lower_bound(b);
  if (it == pool.end() || it->first > b + b / 4) return nullptr;
  void *p = it->second;
  pool.erase(it);
  return p;
}

void *Allocator::alloc(size_t bytes) {
  std::lock_guard<std::mutex> lk(mu_);
  const size_t b = round(bytes);
  if (active_ >= 0) {
    Arena &a = arenas_[active_];
    if (void *p = take(a.free, b)) return p;
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error("device allocation of " + std::to_string(b) +
                      " bytes for a captured graph failed: " + hipGetErrorString(e),
                  -4);
    }
    owner_[p] = {active_, b};
    return p;
  }
  auto it = free_.lower_bound(b);
  if (it != free_.end() && it->first <= b + b / 4) {
    void *p = it->second;
    cached_ -= it->first;
    in_use_ += it->first;
    free_.erase(it);
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // release the cache and retry once
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
    e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error("device allocation of " + std::to_string(b) +
                      " bytes failed: " + hipGetErrorString(e),
                  -4);
    }
  }
  in_use_ += b;
  return p;
}

void Allocator::free(void *p, size_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  const size_t b = round(bytes);
  auto ow = owner_.find(p);
  if (ow != owner_.end()) {
    Arena &a = arenas_[ow->second.first];
    if (defer_)
      a.deferred.emplace_back(b, p);
    else
      a.free.emplace(b, p);
    return;
  }
  in_use_ -= b;
  cached_ += b;
  if (defer_)
    deferred_.emplace_back(b, p);
  else
    free_.emplace(b, p);
}

void Allocator::set_deferred(bool on) {
  std::lock_guard<std::mutex> lk(mu_);
  defer_ = on;
  if (!on) {
    for (auto &kv : deferred_) free_.emplace(kv.first, kv.second);
    deferred_.clear();
    for (auto &ar : arenas_) {
      for (auto &kv : ar.second.deferred) ar.second.free.emplace(kv.first, kv.second);
      ar.second.deferred.clear();
    }
  }
}

int Allocator::arena_create() {
  std::lock_guard<std::mutex> lk(mu_);
  const int id = next_arena_++;
  arenas_[id];
  return id;
}

void Allocator::arena_activate(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  TN_REQUIRE(id < 0 || arenas_.count(id), "arena_activate: no such arena");
  active_ = id;
}

void Allocator::arena_destroy(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = arenas_.find(id);
  if (it == arenas_.end()) return;
  if (active_ == id) active_ = -1;
  // blocks still in use return to the general pool when they are freed
  // (accounted as in use from now on), the free ones join it now
  size_t idle = 0;
  for (auto &kv : it->second.free) {
    free_.emplace(kv.first, kv.second);
    idle += kv.first;
  }
  for (auto &kv : it->second.deferred) {
    deferred_.emplace_back(kv.first, kv.second);
    idle += kv.first;
  }
  size_t all = 0;
  for (auto ow = owner_.begin(); ow != owner_.end();) {
    if (ow->second.first == id) {
      all += ow->second.second;
      ow = owner_.erase(ow);
    } else {
      ++ow;
    }
  }
  cached_ += idle;
  in_use_ += all - idle;
  arenas_.erase(it);
}

void Allocator::release_cached() {
  std::lock_guard<std::mutex> lk(mu_);
  for (auto &kv : free_) (void)hipFree(kv.second);
  cached_ = 0;
  for (auto &kv : deferred_) cached_ += kv.first;
  free_.clear();
}

Allocator::~Allocator() {
  for (auto &kv : free_) (void)hipFree(kv.second);
  for (auto &kv : deferred_) (void)hipFree(kv.second);
  for (auto &ar : arenas_) {
    for (auto &kv : ar.second.free) (void)hipFree(kv.second);
    for (auto &kv : ar.second.deferred) (void)hipFree(kv.second);
  }
}

Storage::~Storage() {
  if (owned && ptr && ctx) ctx->allocator().free(ptr, bytes);
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
Context::Context(int device, hipStream_t stream) : device_(device), alloc_(device) {
  TN_HIP_CHECK(hipSetDevice(device));
  Lane main;
  main.s = stream;
  if (main.s == nullptr) {
    TN_HIP_CHECK(hipStreamCreateWithFlags(&main.s, hipStreamNonBlocking));
    own_stream_ = true;
  }
  kern::linalg_init_device();
  TN_HIP_CHECK(hipMalloc(&main.scal, 4096 * sizeof(double)));
  lanes_.push_back(main);
  TN_HIP_CHECK(hipEventCreateWithFlags(&fork_event_, hipEventDisableTiming));
}

Context::~Context() {
  for (auto &l : lanes_) (void)hipStreamSynchronize(l.s);
  for (auto &kv : sweep_graphs) release_sweep_graph(kv.second);
  sweep_graphs.clear();
  // tensors held by the context return their storage to alloc_, which is
  // destroyed before the public members: release them while it is alive
  svd_start.clear();
  for (auto &p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : event_pool_) (void)hipEventDestroy(e);
  for (auto e : inflight_) (void)hipEventDestroy(e);
  for (auto &kv : plans_) (void)hipFree(kv.second.d_tab);
  for (size_t i = 0; i < lanes_.size(); ++i) {
    if (lanes_[i].ws) alloc_.free(lanes_[i].ws, lanes_[i].ws_bytes);
    if (lanes_[i].scal) (void)hipFree(lanes_[i].scal);
    if (i > 0 || own_stream_) (void)hipStreamDestroy(lanes_[i].s);
  }
  if (fork_event_) (void)hipEventDestroy(fork_event_);
  if (pinned_) (void)hipHostFree(pinned_);
}

void Context::fork(int n) {
  TN_REQUIRE(!forked_, "Context::fork: already forked");
  TN_REQUIRE(n >= 1 && n <= 16, "Context::fork: 1 to 16 lanes");
  TN_REQUIRE(!capturing_ || (int)lanes_.size() >= n + 1,
             "Context::fork: more lanes than prepared for the capture");
  add_lanes(n);
  TN_HIP_CHECK(hipEventRecord(fork_event_, lanes_[0].s));
  for (int i = 1; i <= n; ++i)
    TN_HIP_CHECK(hipStreamWaitEvent(lanes_[i].s, fork_event_, 0));
  alloc_.set_deferred(true);
  nforked_ = n;
  forked_ = true;
  cur_ = 0;
}

void Context::add_lanes(int n) {
  while ((int)lanes_.size() < n + 1) {
    Lane l;
    // default priority: every stream of the library runs at one priority
    TN_HIP_CHECK(hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking));
    TN_HIP_CHECK(hipMalloc(&l.scal, 4096 * sizeof(double)));
    lanes_.push_back(l);
  }
}

void Context::use_lane(int i) {
  TN_REQUIRE(forked_ || i < 0, "Context::use_lane outside a fork");
  TN_REQUIRE(i < nforked_, "Context::use_lane: no such lane");
  cur_ = i < 0 ? 0 : i + 1;
}

void Context::join() {
  TN_REQUIRE(forked_, "Context::join without fork");
  for (int i = 1; i <= nforked_; ++i) {
    hipEvent_t e = get_event();
    TN_HIP_CHECK(hipEventRecord(e, lanes_[i].s));
    TN_HIP_CHECK(hipStreamWaitEvent(lanes_[0].s, e, 0));
    // the wait is enqueued: the event may be recycled once it has completed,
    // so keep it until the next synchronisation point of the main stream
    joined_events_.push_back(e);
  }
  if (joined_events_.size() > 256 && !capturing_) {
    for (auto ev : joined_events_) {
      TN_HIP_CHECK(hipEventSynchronize(ev));
      event_pool_.push_back(ev);
    }
    joined_events_.clear();
  }
  cur_ = 0;
  forked_ = false;
  nforked_ = 0;
  alloc_.set_deferred(false);
}

void Context::join_noexcept() noexcept {
  if (!forked_) return;
  for (int i = 1; i <= nforked_; ++i) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) continue;
    if (hipEventRecord(e, lanes_[i].s) == hipSuccess)
      (void)hipStreamWaitEvent(lanes_[0].s, e, 0);
    (void)hipEventDestroy(e);  // released once the recorded work completes
  }
  (void)hipGetLastError();
  cur_ = 0;
  forked_ = false;
  nforked_ = 0;
  alloc_.set_deferred(false);
}

void Context::move_issued() {
  if (max_moves_ <= 0 || forked_ || capturing_) return;
  hipEvent_t e = get_event();
  TN_HIP_CHECK(hipEventRecord(e, lanes_[0].s));
  inflight_.push_back(e);
  while ((int)inflight_.size() > max_moves_) {
    TN_HIP_CHECK(hipEventSynchronize(inflight_.front()));
    event_pool_.push_back(inflight_.front());
    inflight_.erase(inflight_.begin());
  }
}

void Context::begin_capture() {
  TN_REQUIRE(!capturing_ && !forked_, "begin_capture: already capturing or forked");
  // every lane the captured work may fork into exists before the capture
  add_lanes(max_lanes_);
  // the capture's own scalar areas and (empty) workspaces
  saved_lanes_.clear();
  for (auto &l : lanes_) {
    saved_lanes_.push_back({l.ws, l.ws_bytes, l.scal});
    l.ws = nullptr;
    l.ws_bytes = 0;
    l.scal = nullptr;
  }
  auto restore = [&] {
    for (size_t i = 0; i < lanes_.size(); ++i) {
      if (lanes_[i].scal) alloc_.free(lanes_[i].scal, 4096 * sizeof(double));
      lanes_[i].ws = saved_lanes_[i].ws;
      lanes_[i].ws_bytes = saved_lanes_[i].ws_bytes;
      lanes_[i].scal = saved_lanes_[i].scal;
    }
    saved_lanes_.clear();
  };
  try {
    for (auto &l : lanes_) l.scal = static_cast<double *>(alloc_.alloc(4096 * sizeof(double)));
    TN_HIP_CHECK(hipStreamBeginCapture(lanes_[0].s, hipStreamCaptureModeRelaxed));
  } catch (...) {
    restore();
    throw;
  }
  capturing_ = true;
}

hipGraph_t Context::end_capture(std::vector<std::pair<void *, size_t>> &blocks) {
  TN_REQUIRE(capturing_, "end_capture without begin_capture");
  capturing_ = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(lanes_[0].s, &g);
  blocks.clear();
  for (size_t i = 0; i < lanes_.size(); ++i) {
    Lane &l = lanes_[i];
    if (l.ws) blocks.emplace_back(l.ws, l.ws_bytes);
    blocks.emplace_back(l.scal, 4096 * sizeof(double));
    l.ws = saved_lanes_[i].ws;
    l.ws_bytes = saved_lanes_[i].ws_bytes;
    l.scal = saved_lanes_[i].scal;
  }
  saved_lanes_.clear();
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (g) (void)hipGraphDestroy(g);
    return nullptr;
  }
  return g;
}

int Context::slot_window(int n) {
  const int m = kSlotProjectorSync - kEagerSlots;
  if (slot_used_.empty()) slot_used_.assign(m, false);
  for (int i = 0; i + n <= m; ++i) {
    int j = 0;
    while (j < n && !slot_used_[i + j]) ++j;
    if (j == n) {
      for (j = 0; j < n; ++j) slot_used_[i + j] = true;
      return kEagerSlots + i;
    }
    i += j;  // the slot at i + j is taken: restart after it
  }
  return -1;
}

void Context::release_slot_window(int first, int n) {
  for (int j = 0; j < n; ++j) {
    const int i = first - kEagerSlots + j;
    if (i >= 0 && i < (int)slot_used_.size()) slot_used_[i] = false;
  }
}

void Context::release_sweep_graph(SweepGraph &g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
  for (auto &b : g.blocks) alloc_.free(b.first, b.second);
  g.blocks.clear();
  if (g.arena >= 0) alloc_.arena_destroy(g.arena);
  g.arena = -1;
  if (g.slot_first >= 0) release_slot_window(g.slot_first, g.slot_count);
  g.slot_first = -1;
}

void Context::drop_sweep_graph(int64_t environment) {
  auto it = sweep_graphs.find(environment);
  if (it == sweep_graphs.end()) return;
  // a replay may still run (on any lane): the arena's blocks must not
  // serve other work before it has finished
  for (auto &l : lanes_) (void)hipStreamSynchronize(l.s);
  release_sweep_graph(it->second);
  sweep_graphs.erase(it);
}

void Context::forget_svd_starts(int64_t environment) {
  drop_sweep_graph(environment);
  for (auto it = svd_start.begin(); it != svd_start.end();) {
    if (it->first[0] == environment)
      it = svd_start.erase(it);
    else
      ++it;
  }
}

void *Context::workspace(size_t bytes) {
  Lane &l = lanes_[cur_];
  if (bytes <= l.ws_bytes) return l.ws;
  if (l.ws) alloc_.free(l.ws, l.ws_bytes);
  l.ws_bytes = std::max(bytes, l.ws_bytes * 2);
  l.ws = alloc_.alloc(l.ws_bytes);
  return l.ws;
}

double *Context::scalars(int n) {
  TN_REQUIRE(n <= 4096, "scalar area exhausted");
  return lanes_[cur_].scal;
}

double *Context::pinned(int slot) {
  TN_REQUIRE(slot >= 0 && slot < kPinnedSlots, "pinned slot out of range");
  if (!pinned_)
    TN_HIP_CHECK(hipHostMalloc(&pinned_, (size_t)kPinnedSlots * 512 * sizeof(double)));
  return pinned_ + (size_t)slot * 512;
}

double *Context::pinned_device(int slot) {
  pinned(slot);  // allocates on first use
  if (!pinned_dev_) {
    void *d = nullptr;
    TN_HIP_CHECK(hipHostGetDevicePointer(&d, pinned_, 0));
    pinned_dev_ = static_cast<double *>(d);
  }
  return pinned_dev_ + (size_t)slot * 512;
}

void Context::collect_svd_stats() {
  for (const auto &p : pending_svd) {
    const double *h = pinned(p.slot);
    int rounds = 0;
    std::memcpy(&rounds, h + 511, sizeof(int));
    jacobi_rounds += rounds;
    if (!p.warm) continue;
    auto it = svd_start.find(p.wkey);
    if (it != svd_start.end() && p.nall > 0)
      it->second.range = h[p.nall - 1] > 0.0 ? h[0] / h[p.nall - 1] : 0.0;
  }
  pending_svd.clear();
  eager_slot = 0;
}

void Context::sync() {
  TN_HIP_CHECK(hipStreamSynchronize(stream()));
  if (cur_ == 0 && !joined_events_.empty()) {
    for (auto e : joined_events_) event_pool_.push_back(e);
    joined_events_.clear();
  }
}

hipEvent_t Context::get_event() {
  if (!event_pool_.empty()) {
    hipEvent_t e = event_pool_.back();
    event_pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  TN_HIP_CHECK(hipEventCreate(&e));
  return e;
}

void Context::set_profiling(bool on) {
  if (!on) flush_pending();
  profiling_ = on;
}

void Context::profile_begin(hipEvent_t *a) {
  *a = get_event();
  TN_HIP_CHECK(hipEventRecord(*a, stream()));
}

void Context::profile_end(hipEvent_t a, const char *name, double flops) {
  hipEvent_t b = get_event();
  TN_HIP_CHECK(hipEventRecord(b, stream()));
  pending_.push_back({a, b, name, flops});
  if (pending_.size() >= 4096) flush_pending();
}

void Context::flush_pending() {
  if (pending_.empty()) return;
  for (auto &p : pending_) TN_HIP_CHECK(hipEventSynchronize(p.b));
  for (auto &p : pending_) {
    float ms = 0.f;
    TN_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    auto &st = kstats_[p.name];
    st.launches++;
    st.ms += ms;
    st.flops += p.flops;
    event_pool_.push_back(p.a);
    event_pool_.push_back(p.b);
  }
  pending_.clear();
}

std::map<std::string, Context::KernelStat> Context::kernel_stats() {
  flush_pending();
  return kstats_;
}

void Context::reset_kernel_stats() {
  flush_pending();
  kstats_.clear();
}

std::string shape_str(const Shape &s) {
  std::ostringstream os;
  os << "(";
  for (size_t i = 0; i < s.size(); ++i) os << (i ? "," : "") << s[i];
  os << ")";
  return os.str();
}

static std::vector<int64_t> row_major_strides(const Shape &s) {
  std::vector<int64_t> st(s.size());
  int64_t acc = 1;
  for (int i = (int)s.size() - 1; i >= 0; --i) {
    st[i] = acc;
    acc *= s[i];
  }
  return st;
}

// Offsets of a row-major multi-index over `dims` with the given strides.
static void build_offsets(const std::vector<int64_t> &dims,
                          const std::vector<int64_t> &strides,
                          std::vector<int> &out) {
  int64_t n = 1;
  for (auto d : dims) n *= d;
  out.resize((size_t)n);
  const int r = (int)dims.size();
  std::vector<int64_t> idx(r, 0);
  int64_t off = 0;
  for (int64_t i = 0; i < n; ++i) {
    TN_REQUIRE(off <= INT32_MAX, "tensor too large for 32-bit offsets");
    out[(size_t)i] = (int)off;
    for (int d = r - 1; d >= 0; --d) {
      idx[d]++;
      off += strides[d];
      if (idx[d] < dims[d]) break;
      off -= strides[d] * dims[d];
      idx[d] = 0;
    }
  }
}

static int affine_stride(const std::vector<int> &tab) {
  if (tab.size() <= 1) return 1;
  if (tab[0] != 0) return 0;
  const int s = tab[1];
  if (s <= 0) return 0;
  for (size_t i = 2; i < tab.size(); ++i)
    if ((int64_t)tab[i] != (int64_t)i * s) return 0;
  return s;
}

const ContractPlan &Context::plan(const Shape &sa, const Shape &sb,
                                  const Axes &axa, const Axes &axb,
                                  const Axes &perm) {
  std::ostringstream key;
  key << shape_str(sa) << shape_str(sb) << "|";
  for (auto a : axa) key << a << ",";
  key << "|";
  for (auto a : axb) key << a << ",";
  key << "|";
  for (auto a : perm) key << a << ",";
  auto it = plans_.find(key.str());
  if (it != plans_.end()) return it->second;
  TN_REQUIRE(!capturing_, "tensordot: new contraction plan inside a stream capture");

  const int ra = (int)sa.size(), rb = (int)sb.size(), nc = (int)axa.size();
  TN_REQUIRE((int)axb.size() == nc, "tensordot: axes length mismatch");
  std::vector<int> useda(ra, 0), usedb(rb, 0);
  for (int i = 0; i < nc; ++i) {
    TN_REQUIRE(axa[i] >= 0 && axa[i] < ra && axb[i] >= 0 && axb[i] < rb,
               "tensordot: axis out of range");
    TN_REQUIRE(!useda[axa[i]] && !usedb[axb[i]], "tensordot: repeated axis");
    useda[axa[i]] = usedb[axb[i]] = 1;
    TN_REQUIRE(sa[axa[i]] == sb[axb[i]],
               "tensordot: contracted dimension mismatch " + shape_str(sa) +
                   " x " + shape_str(sb));
  }
  const auto stA = row_major_strides(sa), stB = row_major_strides(sb);
  std::vector<int> freeA, freeB;
  for (int i = 0; i < ra; ++i)
    if (!useda[i]) freeA.push_back(i);
  for (int i = 0; i < rb; ++i)
    if (!usedb[i]) freeB.push_back(i);
  Shape res;
  for (int i : freeA) res.push_back(sa[i]);
  for (int i : freeB) res.push_back(sb[i]);
  const int rr = (int)res.size();
  Axes pm = perm;
  if (pm.empty()) {
    pm.resize(rr);
    std::iota(pm.begin(), pm.end(), 0);
  }
  TN_REQUIRE((int)pm.size() == rr, "tensordot: output permutation rank");
  Shape outs(rr);
  {
    std::vector<int> seen(rr, 0);
    for (int k = 0; k < rr; ++k) {
      TN_REQUIRE(pm[k] >= 0 && pm[k] < rr && !seen[pm[k]],
                 "tensordot: bad output permutation");
      seen[pm[k]] = 1;
      outs[k] = res[pm[k]];
    }
  }
  const auto stO = row_major_strides(outs);
  std::vector<int64_t> cstride(rr);
  for (int k = 0; k < rr; ++k) cstride[pm[k]] = stO[k];

  // Order contracted pairs so that the operand whose unit-stride leg is
  // contracted walks contiguous memory along k.
  std::vector<int> order(nc);
  std::iota(order.begin(), order.end(), 0);
  const bool a_last_contracted = ra > 0 && useda[ra - 1];
  const bool b_last_contracted = rb > 0 && usedb[rb - 1];
  if (a_last_contracted) {
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return stA[axa[x]] > stA[axa[y]];
    });
  } else if (b_last_contracted) {
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return stB[axb[x]] > stB[axb[y]];
    });
  }

  std::vector<int64_t> dm, sm, dk, sak, sbk, dn, sn, cm, cn;
  for (size_t q = 0; q < freeA.size(); ++q) {
    dm.push_back(sa[freeA[q]]);
    sm.push_back(stA[freeA[q]]);
    cm.push_back(cstride[q]);
  }
  for (int q : order) {
    dk.push_back(sa[axa[q]]);
    sak.push_back(stA[axa[q]]);
    sbk.push_back(stB[axb[q]]);
  }
  for (size_t q = 0; q < freeB.size(); ++q) {
    dn.push_back(sb[freeB[q]]);
    sn.push_back(stB[freeB[q]]);
    cn.push_back(cstride[freeA.size() + q]);
  }
  std::vector<int> t_am, t_ak, t_bk, t_bn, t_cm, t_cn;
  build_offsets(dm, sm, t_am);
  build_offsets(dk, sak, t_ak);
  build_offsets(dk, sbk, t_bk);
  build_offsets(dn, sn, t_bn);
  build_offsets(dm, cm, t_cm);
  build_offsets(dn, cn, t_cn);
  // the kernel adds an outer and a contracted offset (and the two output
  // offsets) in 32-bit arithmetic: every such sum must fit as well
  auto maxv = [](const std::vector<int> &t) {
    return t.empty() ? (int64_t)0 : (int64_t)*std::max_element(t.begin(), t.end());
  };
  TN_REQUIRE(maxv(t_am) + maxv(t_ak) <= INT32_MAX &&
                 maxv(t_bk) + maxv(t_bn) <= INT32_MAX &&
                 maxv(t_cm) + maxv(t_cn) <= INT32_MAX,
             "tensordot: operand or result too large for 32-bit offsets " +
                 shape_str(sa) + " x " + shape_str(sb));

  ContractPlan p;
  p.M = (int)t_am.size();
  p.K = (int)t_ak.size();
  p.N = (int)t_bn.size();
  p.out_shape = outs;
  p.a_m_off = 0;
  p.a_k_off = p.a_m_off + p.M;
  p.b_k_off = p.a_k_off + p.K;
  p.b_n_off = p.b_k_off + p.K;
  p.c_m_off = p.b_n_off + p.N;
  p.c_n_off = p.c_m_off + p.M;
  const size_t total = (size_t)p.c_n_off + p.N;
  std::vector<int> all;
  all.reserve(total);
  all.insert(all.end(), t_am.begin(), t_am.end());
  all.insert(all.end(), t_ak.begin(), t_ak.end());
  all.insert(all.end(), t_bk.begin(), t_bk.end());
  all.insert(all.end(), t_bn.begin(), t_bn.end());
  all.insert(all.end(), t_cm.begin(), t_cm.end());
  all.insert(all.end(), t_cn.begin(), t_cn.end());
  TN_HIP_CHECK(hipMalloc(&p.d_tab, std::max<size_t>(total, 1) * sizeof(int)));
  if (total)
    TN_HIP_CHECK(hipMemcpy(p.d_tab, all.data(), total * sizeof(int),
                           hipMemcpyHostToDevice));
  p.a_m_stride = affine_stride(t_am);
  p.a_k_stride = affine_stride(t_ak);
  p.b_k_stride = affine_stride(t_bk);
  p.b_n_stride = affine_stride(t_bn);
  const bool am1 = p.M > 1 && t_am[1] - t_am[0] == 1;
  const bool ak1 = p.K > 1 && t_ak[1] - t_ak[0] == 1;
  p.a_mfast = am1 && !(ak1 && p.K >= 16 && p.M < 16);
  const bool bn1 = p.N > 1 && t_bn[1] - t_bn[0] == 1;
  const bool bk1 = p.K > 1 && t_bk[1] - t_bk[0] == 1;
  p.b_nfast = bn1 && !(bk1 && p.K >= 16 && p.N < 16);
  return plans_.emplace(key.str(), p).first->second;
}

// ---------------------------------------------------------------------------
// Construction
// ---------------------------------------------------------------------------
static Tensor make(Context &ctx, const Shape &shape) {
  Tensor t;
  t.shape = shape;
  auto st = std::make_shared<Storage>();
  st->ctx = &ctx;
  int64_t n = 1;
  for (auto d : shape) {
    TN_REQUIRE(d >= 0, "negative dimension");
    n *= d;
  }
  st->bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(double);
  st->ptr = static_cast<double *>(ctx.allocator().alloc(st->bytes));
  t.st = st;
  return t;
}

Tensor empty(Context &ctx, const Shape &shape) { return make(ctx, shape); }

Tensor zeros(Context &ctx, const Shape &shape) {
  Tensor t = make(ctx, shape);
  kern::fill(t.data(), t.size(), 0.0, ctx.stream());
  return t;
}

Tensor from_host(Context &ctx, const Shape &shape, const double *host) {
  Tensor t = make(ctx, shape);
  if (t.size())
    TN_HIP_CHECK(hipMemcpyAsync(t.data(), host, t.size() * sizeof(double),
                                hipMemcpyHostToDevice, ctx.stream()));
  ctx.sync();
  return t;
}

void to_host(const Tensor &t, double *host) {
  TN_REQUIRE(t.defined(), "to_host: undefined tensor");
  Context &ctx = *t.ctx();
  if (t.size())
    TN_HIP_CHECK(hipMemcpyAsync(host, t.data(), t.size() * sizeof(double),
                                hipMemcpyDeviceToHost, ctx.stream()));
  ctx.sync();
}

Tensor borrow(Context &ctx, const Shape &shape, double *dev) {
  Tensor t;
  t.shape = shape;
  auto st = std::make_shared<Storage>();
  st->ctx = &ctx;
  st->ptr = dev;
  st->owned = false;
  int64_t n = 1;
  for (auto d : shape) n *= d;
  st->bytes = (size_t)n * sizeof(double);
  t.st = st;
  return t;
}

Tensor clone(const Tensor &t) {
  Context &ctx = *t.ctx();
  Tensor r = make(ctx, t.shape);
  if (t.size())
    TN_HIP_CHECK(hipMemcpyAsync(r.data(), t.data(), t.size() * sizeof(double),
                                hipMemcpyDeviceToDevice, ctx.stream()));
  return r;
}

// ---------------------------------------------------------------------------
// Shape operations
// ---------------------------------------------------------------------------
Tensor reshape(const Tensor &t, const Shape &shape) {
  int64_t n = 1;
  for (auto d : shape) n *= d;
  TN_REQUIRE(n == t.size(), "reshape: size mismatch " + shape_str(t.shape) +
                                " -> " + shape_str(shape));
  Tensor r = t;
  r.shape = shape;
  return r;
}

Tensor transpose(const Tensor &t, const Axes &axes) {
  const int r = t.rank();
  TN_REQUIRE((int)axes.size() == r, "transpose: rank mismatch");
  Context &ctx = *t.ctx();
  const auto st = row_major_strides(t.shape);
  Shape ns(r);
  std::vector<int64_t> src_str(r);
  for (int k = 0; k < r; ++k) {
    TN_REQUIRE(axes[k] >= 0 && axes[k] < r, "transpose: bad axis");
    ns[k] = t.shape[axes[k]];
    src_str[k] = st[axes[k]];
  }
  Tensor out = make(ctx, ns);
  const auto dst_str = row_major_strides(ns);
  kern::strided_copy(t.data(), out.data(), r, ns.data(), src_str.data(),
                     dst_str.data(), out.size(), ctx.stream());
  return out;
}

void tensordot_into(const Tensor &A, const Tensor &B, const Axes &axA,
                    const Axes &axB, const Axes &perm, Tensor &C, double alpha,
                    double beta) {
  Context &ctx = *A.ctx();
  const ContractPlan &p = ctx.plan(A.shape, B.shape, axA, axB, perm);
  TN_REQUIRE(C.shape == p.out_shape, "tensordot_into: output shape mismatch");
  kern::GemmArgs g;
  g.A = A.data();
  g.B = B.data();
  g.C = C.data();
  g.a_m = p.d_tab + p.a_m_off;
  g.a_k = p.d_tab + p.a_k_off;
  g.b_k = p.d_tab + p.b_k_off;
  g.b_n = p.d_tab + p.b_n_off;
  g.c_m = p.d_tab + p.c_m_off;
  g.c_n = p.d_tab + p.c_n_off;
  g.M = p.M;
  g.N = p.N;
  g.K = p.K;
  g.alpha = alpha;
  g.beta = beta;
  g.a_mfast = p.a_mfast;
  g.b_nfast = p.b_nfast;
  g.a_m_stride = p.a_m_stride;
  g.a_k_stride = p.a_k_stride;
  g.b_k_stride = p.b_k_stride;
  g.b_n_stride = p.b_n_stride;
  g.fp32 = ctx.contraction_mode() == 1;
  g.bf16x3 = ctx.contraction_mode() == 2;
  const double flops = 2.0 * p.M * (double)p.N * p.K;
  // autotuning (C = A B overwrites C, so every candidate leaves a valid
  // result): time each launch shape twice on this stream, keep the fastest
  if (ctx.autotune() && !ctx.capturing() && beta == 0.0 && p.tuned_bm == 0 &&
      flops >= Context::kAutotuneFlops) {
    hipEvent_t e0, e1;
    TN_HIP_CHECK(hipEventCreate(&e0));
    TN_HIP_CHECK(hipEventCreate(&e1));
    float best = 0.0f;
    for (const kern::GemmPlan &c : kern::gemm_candidates(p.M, p.N, p.K)) {
      const size_t cws = kern::gemm_workspace_bytes(p.M, p.N, p.K, &c);
      void *cw = cws ? ctx.workspace(cws) : nullptr;
      TN_HIP_CHECK(hipEventRecord(e0, ctx.stream()));
      kern::gemm_f64(g, cw, cws, ctx.stream(), &c);
      kern::gemm_f64(g, cw, cws, ctx.stream(), &c);
      TN_HIP_CHECK(hipEventRecord(e1, ctx.stream()));
      TN_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      TN_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (p.tuned_bm == 0 || ms < best) {
        best = ms;
        p.tuned_bm = c.bm;
        p.tuned_splits = c.splits;
      }
    }
    TN_HIP_CHECK(hipEventDestroy(e0));
    TN_HIP_CHECK(hipEventDestroy(e1));
  }
  kern::GemmPlan tuned{p.tuned_bm, p.tuned_splits};
  const kern::GemmPlan *plan = p.tuned_bm ? &tuned : nullptr;
  static const kern::GemmPlan kStreamK{128, 0};
  if (!plan && ctx.stream_k() && !g.fp32 && !g.bf16x3 && kern::gemm_stream_k(p.M, p.N, p.K))
    plan = &kStreamK;
  const size_t ws = kern::gemm_workspace_bytes(p.M, p.N, p.K, plan);
  void *w = ws ? ctx.workspace(ws) : nullptr;
  if (ctx.profiling()) {
    hipEvent_t a;
    ctx.profile_begin(&a);
    const char *name = kern::gemm_f64(g, w, ws, ctx.stream(), plan);
    // one report line per instantiation and problem shape
    const std::string key = std::string(name) + "[" + std::to_string(p.M) +
                            "x" + std::to_string(p.N) + "x" +
                            std::to_string(p.K) + "]";
    ctx.profile_end(a, key.c_str(), flops);
  } else {
    kern::gemm_f64(g, w, ws, ctx.stream(), plan);
  }
  ctx.gemm_flops += flops;
  ctx.gemm_calls++;
}

Tensor tensordot_t(const Tensor &A, const Tensor &B, const Axes &axA,
                   const Axes &axB, const Axes &perm) {
  TN_REQUIRE(A.defined() && B.defined(), "tensordot: undefined operand");
  Context &ctx = *A.ctx();
  const ContractPlan &p = ctx.plan(A.shape, B.shape, axA, axB, perm);
  Tensor C = make(ctx, p.out_shape);
  tensordot_into(A, B, axA, axB, perm, C, 1.0, 0.0);
  return C;
}

Tensor tensordot(const Tensor &A, const Tensor &B, const Axes &axA,
                 const Axes &axB) {
  return tensordot_t(A, B, axA, axB, Axes{});
}

Tensor slice(const Tensor &t, int axis, int64_t begin, int64_t end) {
  const int r = t.rank();
  TN_REQUIRE(axis >= 0 && axis < r, "slice: bad axis");
  TN_REQUIRE(0 <= begin && begin <= end && end <= t.shape[axis],
             "slice: bad range");
  Context &ctx = *t.ctx();
  Shape ns = t.shape;
  ns[axis] = end - begin;
  Tensor out = make(ctx, ns);
  const auto st = row_major_strides(t.shape);
  const auto dst = row_major_strides(ns);
  const double *src = t.data() + begin * st[axis];
  kern::strided_copy(src, out.data(), r, ns.data(), st.data(), dst.data(),
                     out.size(), ctx.stream());
  return out;
}

Tensor extend(const Tensor &t, const Shape &shape) {
  const int r = t.rank();
  TN_REQUIRE((int)shape.size() == r, "extend: rank mismatch");
  for (int i = 0; i < r; ++i)
    TN_REQUIRE(shape[i] >= t.shape[i], "extend: target smaller than source");
  Context &ctx = *t.ctx();
  Tensor out = zeros(ctx, shape);
  const auto st = row_major_strides(t.shape);
  const auto dst = row_major_strides(shape);
  kern::strided_copy(t.data(), out.data(), r, t.shape.data(), st.data(),
                     dst.data(), t.size(), ctx.stream());
  return out;
}

Tensor conj(const Tensor &t) { return t; }

Tensor bra(const Tensor &t) {
  Context &ctx = *t.ctx();
  if (!ctx.single_layer()) return conj(t);
  TN_REQUIRE(t.rank() == 5 && t.shape[4] == 1,
             "single-layer site tensors must be (l, u, r, d, 1)");
  Tensor u = make(ctx, Shape{1, 1, 1, 1, 1});
  kern::fill(u.data(), 1, 1.0, ctx.stream());
  return u;
}

void trace_dev(const Tensor &A, const Tensor &B, const Axes &axA,
               const Axes &axB, double *out) {
  TN_REQUIRE(A.rank() == (int)axA.size() && B.rank() == (int)axB.size() &&
                 axA.size() == axB.size(),
             "trace: every leg must be contracted");
  Context &ctx = *A.ctx();
  // Bring A into B's leg order: A' leg axB[k] = A leg axA[k].
  const int r = A.rank();
  Axes perm(r);
  for (int k = 0; k < r; ++k) perm[axB[k]] = axA[k];
  bool ident = true;
  for (int k = 0; k < r; ++k) ident = ident && perm[k] == k;
  Tensor Ap = ident ? A : transpose(A, perm);
  TN_REQUIRE(Ap.shape == B.shape, "trace: shape mismatch");
  const int np = kern::dot_partials(B.size());
  double *part = static_cast<double *>(ctx.workspace(np * sizeof(double)));
  kern::dot(Ap.data(), B.data(), B.size(), out, part, ctx.stream());
}

double trace(const Tensor &A, const Tensor &B, const Axes &axA,
             const Axes &axB) {
  Context &ctx = *A.ctx();
  double *d = ctx.scalars(1);
  trace_dev(A, B, axA, axB, d);
  double h = 0.0;
  TN_HIP_CHECK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost,
                              ctx.stream()));
  ctx.sync();
  return h;
}

double max_abs(const Tensor &t) {
  Context &ctx = *t.ctx();
  double *d = ctx.scalars(1);
  const int np = kern::max_abs_partials(t.size());
  double *part = static_cast<double *>(ctx.workspace(np * sizeof(double)));
  kern::max_abs(t.data(), t.size(), d, part, ctx.stream());
  double h = 0.0;
  TN_HIP_CHECK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost,
                              ctx.stream()));
  ctx.sync();
  return h;
}

void normalize_max_abs(Tensor &t) {
  Context &ctx = *t.ctx();
  double *d = ctx.scalars(1);
  const int np = kern::max_abs_partials(t.size());
  double *part = static_cast<double *>(ctx.workspace(np * sizeof(double)));
  kern::max_abs(t.data(), t.size(), d, part, ctx.stream());
  kern::scale_by_inv(t.data(), t.size(), d, ctx.stream());
}

void multiply_vector_dev(Tensor &t, const double *v_dev, int axis) {
  TN_REQUIRE(axis >= 0 && axis < t.rank(), "multiply_vector: bad axis");
  int64_t inner = 1;
  for (int i = axis + 1; i < t.rank(); ++i) inner *= t.shape[i];
  kern::scale_axis(t.data(), t.size(), v_dev, t.shape[axis], inner,
                   t.ctx()->stream());
}

void scale_inplace(Tensor &t, double a) {
  kern::scale(t.data(), t.size(), a, t.ctx()->stream());
}

Tensor add(const Tensor &a, const Tensor &b, double alpha, double beta) {
  TN_REQUIRE(a.shape == b.shape, "add: shape mismatch");
  Tensor r = clone(a);
  if (alpha != 1.0) scale_inplace(r, alpha);
  kern::axpy(r.data(), b.data(), r.size(), beta, r.ctx()->stream());
  return r;
}

}  // namespace tn
