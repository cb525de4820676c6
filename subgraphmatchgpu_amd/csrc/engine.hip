// engine.hip -- context, resident graph and the level/batch engine (engine.h).
//
// Host orchestration only; every device step is a kernel from kernels.hip.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <memory>
#include <new>
#include <cstring>
#include <utility>

#include "graph.h"

namespace sme {

void Fail(Status s, const std::string& msg) { throw Error{s, msg}; }

#define SME_HIP(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) Fail(kErrHip, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

const char* SlotName(int slot) {
  static const char* kNames[kNumSlots] = {"prep",  "scan",   "partition", "intersect",
                                          "intersect_lane", "filter", "count", "compact",
                                          "expand", "intersect_group", "clique_small",
                                          "clique_rows", "clique_count", "house_mirror",
                                          "house_wedge", "misc"};
  return slot >= 0 && slot < kNumSlots ? kNames[slot] : "?";
}

// ---------------------------------------------------------------------------
Context::~Context() {
  for (auto& m : marks) {
    (void)hipEventDestroy(m.start);
    (void)hipEventDestroy(m.stop);
  }
  for (auto e : spare_events) (void)hipEventDestroy(e);
  own_pool_.Release();
  if (pinned_) (void)hipHostFree(pinned_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

Status Context::Open(int device, uint64_t budget_bytes, std::string* err) {
  const Status s = OpenShared(device, nullptr, err);
  if (s != kOk) return s;
  if (!own_pool_.Init(budget_bytes)) {
    *err = "cannot reserve the device pool of " + std::to_string(budget_bytes) + " bytes";
    return kErrOom;
  }
  return kOk;
}

Status Context::OpenShared(int device, DevicePool* shared, std::string* err) {
  pool_ = shared ? shared : &own_pool_;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    *err = "no HIP device is visible";
    return kErrHip;
  }
  if (device < 0 || device >= count) {
    *err = "device id out of range";
    return kErrInvalid;
  }
  device_ = device;
  if (hipSetDevice(device) != hipSuccess) {
    *err = "hipSetDevice failed";
    return kErrHip;
  }
  // default-priority stream
  if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) {
    *err = "hipStreamCreate failed";
    return kErrHip;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&pinned_), 4096, hipHostMallocDefault) != hipSuccess) {
    *err = "hipHostMalloc failed";
    return kErrHip;
  }
  // the no-allocation entry point's scratch; a failure is reported by that
  // entry point, not here
  std::string scratch_err;
  (void)InitStaticScratch(device, &scratch_err);
  return kOk;
}

void* Context::Alloc(size_t bytes) {
  void* p = pool_->Alloc(bytes, stream_);
  if (p) {
    held += DevicePool::Rounded(bytes);
    if (held > held_peak) held_peak = held;
  }
  return p;
}

void Context::Free(void* p, size_t bytes) {
  if (!p) return;
  pool_->Free(p, stream_);
  held -= DevicePool::Rounded(bytes);
}

uint64_t Context::RunAvailable() const {
  uint64_t avail = std::min<uint64_t>(pool_->Available(), pool_->LargestFree());
  if (limit) avail = std::min<uint64_t>(avail, limit > held ? limit - held : 0);
  return avail;
}

hipEvent_t Context::TakeEvent() {
  if (!spare_events.empty()) {
    hipEvent_t e = spare_events.back();
    spare_events.pop_back();
    return e;
  }
  hipEvent_t e;
  SME_HIP(hipEventCreate(&e));
  return e;
}

void Context::ResolveMarks() {
  for (auto& m : marks) {
    float ms = 0;
    SME_HIP(hipEventElapsedTime(&ms, m.start, m.stop));
    SlotStats& s = slots[m.slot];
    s.launches++;
    s.ms += ms;
    s.bytes += m.bytes;
    s.items += m.items;
    spare_events.push_back(m.start);
    spare_events.push_back(m.stop);
  }
  marks.clear();
}

// ---------------------------------------------------------------------------
DevGraph::~DevGraph() {
  if (device >= 0) (void)hipSetDevice(device);
  if (rp) (void)hipFree(rp);
  if (cols) (void)hipFree(cols);
  if (hub_id) (void)hipFree(hub_id);
  if (hub_bits) (void)hipFree(hub_bits);
  if (up) (void)hipFree(up);
  if (upper.dverts) (void)hipFree(upper.dverts);
  if (upper.drf) (void)hipFree(upper.drf);
  if (upper.dwf) (void)hipFree(upper.dwf);
  if (upper.dcf) (void)hipFree(upper.dcf);
  delete ranked;
}

Status DevGraph::Upload(Context* ctx, const uint64_t* row_ptrs, const uint32_t* c, uint64_t nv,
                        uint64_t ne, std::string* err) {
  bool loops = false;
  if (!row_ptrs || (ne && !c)) {
    *err = "null graph arrays";
    return kErrInvalid;
  }
  if (nv >= 0xffffffffull) {
    *err = "vertex ids must fit in 32 bits";
    return kErrInvalid;
  }
  if (!CheckCsr(nv, ne, row_ptrs, c, &loops, err)) return kErrInvalid;
  device = ctx->device();
  SME_HIP(hipSetDevice(device));
  n = nv;
  nnz = ne;
  self_loops = loops;
  for (uint64_t v = 0; v < nv; ++v) {
    const uint32_t* b = c + row_ptrs[v];
    const uint32_t* e = c + row_ptrs[v + 1];
    const uint64_t dp = e - std::upper_bound(b, e, static_cast<uint32_t>(v));
    max_degree = std::max<uint64_t>(max_degree, e - b);
    max_up_degree = std::max<uint64_t>(max_up_degree, dp);
    if (dp > kLocalMax) {
      upper.huge.push_back(static_cast<uint32_t>(v));
    } else if (dp > kLocalSmall) {
      upper.verts.push_back(static_cast<uint32_t>(v));
      upper.d.push_back(static_cast<uint32_t>(dp));
    }
  }
  {
    const uint64_t nb = upper.verts.size();
    std::vector<uint32_t> order(nb);
    for (uint64_t x = 0; x < nb; ++x) order[x] = static_cast<uint32_t>(x);
    std::sort(order.begin(), order.end(), [&](uint32_t p, uint32_t q) {
      return upper.d[p] != upper.d[q] ? upper.d[p] > upper.d[q] : upper.verts[p] < upper.verts[q];
    });
    std::vector<uint32_t> sv(nb), sd(nb);
    upper.rf.assign(nb + 1, 0);
    upper.wf.assign(nb + 1, 0);
    upper.cf.assign(nb + 1, 0);
    for (uint64_t x = 0; x < nb; ++x) {
      sv[x] = upper.verts[order[x]];
      sd[x] = upper.d[order[x]];
      const uint64_t d = sd[x];
      upper.rf[x + 1] = upper.rf[x] + d;
      upper.wf[x + 1] = upper.wf[x] + d * ((d + 63) / 64);
      upper.cf[x + 1] = upper.cf[x] + (d + 63) / 64;
    }
    upper.verts.swap(sv);
    upper.d.swap(sd);
    if (nb) {
      if (hipMalloc(reinterpret_cast<void**>(&upper.dverts), nb * sizeof(uint32_t)) != hipSuccess ||
          hipMalloc(reinterpret_cast<void**>(&upper.drf), (nb + 1) * sizeof(uint64_t)) != hipSuccess ||
          hipMalloc(reinterpret_cast<void**>(&upper.dwf), (nb + 1) * sizeof(uint64_t)) != hipSuccess ||
          hipMalloc(reinterpret_cast<void**>(&upper.dcf), (nb + 1) * sizeof(uint64_t)) != hipSuccess) {
        (void)hipGetLastError();
        *err = "hipMalloc of the upper-neighbourhood lists failed";
        return kErrOom;
      }
      SME_HIP(hipMemcpyAsync(upper.dverts, upper.verts.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice,
                             ctx->stream()));
      SME_HIP(hipMemcpyAsync(upper.drf, upper.rf.data(), (nb + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                             ctx->stream()));
      SME_HIP(hipMemcpyAsync(upper.dwf, upper.wf.data(), (nb + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                             ctx->stream()));
      SME_HIP(hipMemcpyAsync(upper.dcf, upper.cf.data(), (nb + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                             ctx->stream()));
    }
  }
  if (hipMalloc(reinterpret_cast<void**>(&rp), (nv + 1) * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&cols), (ne ? ne : 1) * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    *err = "hipMalloc of the CSR failed";
    return kErrOom;
  }
  SME_HIP(hipMemcpyAsync(rp, row_ptrs, (nv + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream()));
  if (ne)
    SME_HIP(hipMemcpyAsync(cols, c, ne * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream()));
  SME_HIP(hipStreamSynchronize(ctx->stream()));
  if (ctx->hub_min_degree && ctx->hub_max_bytes && nv) {
    // highest-degree vertices first, as many rows as the byte cap allows
    hub_words = (nv + 31) / 32;
    std::vector<uint32_t> cand;
    for (uint64_t v = 0; v < nv; ++v)
      if (row_ptrs[v + 1] - row_ptrs[v] >= ctx->hub_min_degree) cand.push_back(static_cast<uint32_t>(v));
    std::sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) {
      const uint64_t da = row_ptrs[a + 1] - row_ptrs[a], db = row_ptrs[b + 1] - row_ptrs[b];
      return da != db ? da > db : a < b;
    });
    const uint64_t row_bytes = hub_words * sizeof(uint32_t);
    nhub = std::min<uint64_t>(cand.size(), ctx->hub_max_bytes / row_bytes);
    cand.resize(nhub);
    if (nhub) {
      std::vector<uint32_t> ids(nv, kNoHub);
      for (uint64_t h = 0; h < nhub; ++h) ids[cand[h]] = static_cast<uint32_t>(h);
      uint32_t* hubs = nullptr;
      if (hipMalloc(reinterpret_cast<void**>(&hub_id), nv * sizeof(uint32_t)) != hipSuccess ||
          hipMalloc(reinterpret_cast<void**>(&hub_bits), nhub * row_bytes) != hipSuccess ||
          hipMalloc(reinterpret_cast<void**>(&hubs), nhub * sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        if (hubs) (void)hipFree(hubs);
        *err = "hipMalloc of the hub bitmaps failed";
        return kErrOom;
      }
      SME_HIP(hipMemcpyAsync(hub_id, ids.data(), nv * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream()));
      SME_HIP(hipMemcpyAsync(hubs, cand.data(), nhub * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream()));
      LaunchFill32(hub_bits, nhub * hub_words, 0, ctx->stream());
      LaunchBuildHubBits(rp, cols, hubs, nhub, hub_words, hub_bits, ctx->stream());
      SME_HIP(hipStreamSynchronize(ctx->stream()));
      (void)hipFree(hubs);
    }
  }
  if (hipMalloc(reinterpret_cast<void**>(&up), (nv ? nv : 1) * sizeof(uint64_t)) != hipSuccess) {
    (void)hipGetLastError();
    *err = "hipMalloc of the list split points failed";
    return kErrOom;
  }
  LaunchUpperStart(rp, cols, nv, up, ctx->stream());
  SME_HIP(hipStreamSynchronize(ctx->stream()));
  if (ctx->rank_order && nv) {
    std::vector<uint64_t> rrp;
    std::vector<uint32_t> rcols, new_id;
    RankByDegree(nv, row_ptrs, c, &rrp, &rcols, &new_id);
    ranked = new DevGraph;
    // the copy has no copy of its own (restored on every exit)
    struct Restore {
      Context* c;
      ~Restore() { c->rank_order = true; }
    } restore{ctx};
    ctx->rank_order = false;
    const Status s = ranked->Upload(ctx, rrp.data(), rcols.data(), nv, ne, err);
    if (s != kOk) return s;
  }
  return kOk;
}

// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kMaxBatchItems = 1ull << 31;
constexpr uint64_t kMaxBatchPaths = 1ull << 31;

// Pool-backed device array, released when it leaves scope.
template <typename T>
class Buf {
 public:
  Buf(Context* ctx, uint64_t n) : ctx_(ctx), n_(n) {
    if (n == 0) return;
    p_ = static_cast<T*>(ctx->Alloc(n * sizeof(T)));
    if (!p_)
      Fail(kErrOom, "device pool exhausted: need " + std::to_string(n * sizeof(T)) + " bytes, " +
                        std::to_string(ctx->pool().Available()) + " available");
  }
  ~Buf() { Reset(); }
  void Reset() {
    if (p_) ctx_->Free(p_, n_ * sizeof(T));
    p_ = nullptr;
  }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  T* get() const { return p_; }
  uint64_t size() const { return n_; }

 private:
  Context* ctx_;
  uint64_t n_;
  T* p_ = nullptr;
};

struct Frame {
  uint64_t P = 0;
  ColSet cs{};
  uint32_t matmask = 0;
  CandRef cand[kQ]{};
  uint32_t candmask = 0;
  // candidate segment of the vertex materialized last, per path (for reuse)
  CandRef reuse{};
  bool has_reuse = false;
  // cand[excl_u] is a view that still holds the values M[excl_v[k]] marked
  // per path in excl_in (a FILTER_COMPUTE of != conditions left uncopied)
  int excl_u = -1;
  uint8_t nexcl = 0;
  uint8_t excl_v[2]{};
  const uint8_t* excl_in = nullptr;
};

ColSet SliceCols(const ColSet& c, uint64_t l) {
  ColSet r = c;
  for (int w = 0; w < kQ; ++w) {
    if (r.inst[w])
      r.inst[w] += l;
    else
      r.iota[w] += static_cast<uint32_t>(l);
  }
  return r;
}

CandRef SliceCand(const CandRef& c, uint64_t l) {
  CandRef r = c;
  if (r.idx) {
    r.idx += l;
  } else {
    if (r.beg) r.beg += l;
    if (r.len) r.len += l;
  }
  return r;
}

Frame SliceFrame(const Frame& f, uint64_t l, uint64_t r) {
  Frame s = f;
  s.P = r - l;
  s.cs = SliceCols(f.cs, l);
  for (int w = 0; w < kQ; ++w)
    if (f.candmask & (1u << w)) s.cand[w] = SliceCand(f.cand[w], l);
  if (f.has_reuse) s.reuse = SliceCand(f.reuse, l);
  if (f.excl_in) s.excl_in = f.excl_in + l;
  return s;
}

// Timing scope around one launch.
class Timed {
 public:
  Timed(Context* c, hipStream_t st, int slot, double bytes, uint64_t items)
      : c_(c), st_(st), slot_(slot), bytes_(bytes), items_(items) {
    if (c_->profiling) {
      start_ = c_->TakeEvent();
      SME_HIP(hipEventRecord(start_, st_));
    }
  }
  ~Timed() {
    if (!c_->profiling || !start_) return;
    hipEvent_t stop = nullptr;
    if (!c_->spare_events.empty()) {
      stop = c_->spare_events.back();
      c_->spare_events.pop_back();
    } else if (hipEventCreate(&stop) != hipSuccess) {
      c_->spare_events.push_back(start_);
      return;
    }
    if (hipEventRecord(stop, st_) != hipSuccess) {
      c_->spare_events.push_back(stop);
      c_->spare_events.push_back(start_);
      return;
    }
    c_->marks.push_back({start_, stop, slot_, bytes_, items_});
  }

 private:
  Context* c_;
  hipStream_t st_;
  int slot_;
  double bytes_;
  uint64_t items_;
  hipEvent_t start_ = nullptr;
};

class Runner {
 public:
  Runner(Context* ctx, const DevGraph* g, const EnginePlan* plan)
      : ctx_(ctx), g_(g), plan_(plan), st_(ctx->stream()), pool_(ctx) {}

  // Match-list output (full-materialization plans): the tuples of the last
  // level are appended to host_out (row-major, plan vertex order) up to cap
  // rows, in the order the batches finish, which is lexicographic.
  void SetMatchOutput(uint32_t* host_out, uint64_t cap) {
    list_ = true;
    list_out_ = host_out;
    list_cap_ = cap;
    list_rows_ = 0;
  }
  uint64_t ListedRows() const { return list_rows_; }

  // Per-bucket counts (device arrays owned by the caller).
  void SetBuckets(const uint32_t* dev_bounds, uint64_t nbuckets, unsigned long long* dev_counts,
                  const uint32_t* dev_table, uint32_t base) {
    bk_.bounds = dev_bounds;
    bk_.nbuckets = nbuckets;
    bk_.counts = dev_counts;
    bk_.start_v = static_cast<uint8_t>(plan_->ops[0].u);
    bk_.table = dev_table;
    bk_.base = base;
  }

  void Run(uint32_t v_begin, uint32_t v_end, uint64_t* count, uint64_t* levels, RunStats* stats) {
    Analyze();
    // [total, per-level counts, per-level items counted by lanes, per-level
    // values streamed by group counting, per-level group walk lower bound,
    // per-level items walked by workgroup-per-path counting]
    const int nacc = 1 + 5 * plan_->nops;
    Buf<unsigned long long> acc(pool_, nacc);
    acc_ = acc.get();
    // a fill kernel, not hipMemsetAsync: callers may capture the stream
    LaunchFill64(reinterpret_cast<uint64_t*>(acc_), nacc, 0, st_);
    std::memset(lane_lists_, 0, sizeof(lane_lists_));
    std::memset(wg_lists_, 0, sizeof(wg_lists_));
    std::memset(host_level_, 0, sizeof(host_level_));
    host_count_ = 0;
    if (v_end > v_begin) Level0(v_begin, v_end);
    unsigned long long dev[1 + 5 * kMaxOps];
    SME_HIP(hipMemcpyAsync(ctx_->pinned(), acc_, sizeof(unsigned long long) * nacc,
                           hipMemcpyDeviceToHost, st_));
    SME_HIP(hipStreamSynchronize(st_));
    std::memcpy(dev, ctx_->pinned(), sizeof(unsigned long long) * nacc);
    stats_.d2h_syncs++;
    if (ctx_->profiling) {
      // lane-counted items are known only now: 4 bytes per list per item
      SlotStats& ls = ctx_->slots[kSlotLane];
      for (int L = 0; L < plan_->nops; ++L) {
        const uint64_t items = dev[1 + plan_->nops + L];
        ls.bytes += 4.0 * lane_lists_[L] * static_cast<double>(items);
        ls.items += items;
      }
      // group counting: 4 bytes per list (two) per item of the shorter
      // window, the walk of a pivot-and-probe intersection
      SlotStats& gs = ctx_->slots[kSlotGroup];
      for (int L = 0; L < plan_->nops; ++L) {
        gs.items += dev[1 + 2 * plan_->nops + L];
        gs.bytes += 8.0 * static_cast<double>(dev[1 + 3 * plan_->nops + L]);
      }
      // workgroup-per-path counting: 4 bytes per list per pivot item, as the
      // item schedule of the same level
      SlotStats& is = ctx_->slots[kSlotIntersect];
      for (int L = 0; L < plan_->nops; ++L) {
        const uint64_t items = dev[1 + 4 * plan_->nops + L];
        is.items += items;
        is.bytes += 4.0 * wg_lists_[L] * static_cast<double>(items);
      }
    }
    uint64_t total = host_count_ + dev[0];
    for (int L = 0; L < plan_->nops; ++L) levels_[L] = host_level_[L] + dev[1 + L];
    // GPSM counts the size of the last level (HybridGPSMGPUProcessor.cuh:23-33)
    if (plan_->algo == kAlgoGpsm || plan_->algo == kAlgoGpsmReuse) total = levels_[plan_->nops - 1];
    *count = total;
    if (levels)
      for (int L = 0; L < plan_->nops; ++L) levels[L] = levels_[L];
    if (stats) *stats = stats_;
  }

 private:
  // ----- plan analysis -----
  void Analyze() {
    const int n = plan_->nops;
    uint32_t need_after = 0;
    for (int L = n - 1; L >= 0; --L) {
      live_after_[L] = need_after;
      const OpDesc& op = plan_->ops[L];
      uint32_t ref = 0;
      for (int i = 0; i < op.nc; ++i) ref |= 1u << op.cond[i].v;
      if (op.kind == kCompute || op.kind == kComputePathCount || op.kind == kComputeCount)
        for (int i = 0; i < op.nb; ++i) ref |= 1u << op.bconn[i];
      // match lists read every column at the last level
      if (list_) live_after_[L] |= (1u << plan_->nv) - 1;
      // per-bucket counting reads the first matched vertex at every level
      if (bk_.bounds) {
        ref |= 1u << bk_.start_v;
        live_after_[L] |= 1u << bk_.start_v;
      }
      need_after |= ref;
    }
    std::memset(compute_bconn_, 0, sizeof(compute_bconn_));
    int compute_level[kQ];
    for (int w = 0; w < kQ; ++w) compute_level[w] = -1;
    for (int L = 0; L < n; ++L) {
      const OpDesc& op = plan_->ops[L];
      if ((op.kind == kCompute || op.kind == kComputePathCount) && !compute_bconn_[op.u]) {
        for (int i = 0; i < op.nb; ++i) compute_bconn_[op.u] |= 1u << op.bconn[i];
        compute_level[op.u] = L;
      }
    }
    // Intersection reuse.  Level L computing u may start from the candidate
    // segment of w (materialized at level L-1) when
    //   - w's set was produced by COMPUTE (values kept, no FILTER since),
    //   - bconn(w) + {w} is contained in bconn(u),
    //   - every condition that filtered w's set is implied for u's set:
    //     non-equality always is (u's compute condition lists every vertex
    //     materialized before it), and x <op> M[v] holds when u has the same
    //     condition on v, or on w (then M[w] <op> M[v] gives it).
    // The reused set then equals the intersection of bconn(w) lists passing
    // u's conditions, so u's candidates are unchanged.
    for (int L = 0; L < n; ++L) {
      reuse_w_[L] = -1;
      keep_reuse_[L] = false;
      nrem_[L] = 0;
      alias_of_[L] = -1;
    }
    // Identical intersections in one frame (e.g. the diamond's CPC1 and
    // CPC3, both N(v0) and N(v2) with only implied non-equalities): the
    // later level takes the earlier level's candidate set.
    for (int L = 1; L < n; ++L) {
      const OpDesc& op = plan_->ops[L];
      if (op.kind != kCompute && op.kind != kComputePathCount) continue;
      for (int Lp = L - 1; Lp >= 1; --Lp) {
        const OpDesc& pr = plan_->ops[Lp];
        if (pr.kind != kCompute && pr.kind != kComputePathCount) break;  // frame changed
        if (op.kind == kCompute && pr.kind != kCompute) continue;       // values needed
        if (SameIntersection(op, pr)) {
          alias_of_[L] = Lp;
          break;
        }
      }
    }
    if (!plan_->reuse) return;
    for (int L = 2; L < n; ++L) {
      const OpDesc& op = plan_->ops[L];
      if (op.kind != kCompute && op.kind != kComputePathCount && op.kind != kComputeCount) continue;
      const OpDesc& prev = plan_->ops[L - 1];
      if (prev.kind != kMaterialize) continue;
      const int w = prev.u;
      const int Lw = compute_level[w];
      if (Lw < 0 || plan_->ops[Lw].kind != kCompute) continue;
      bool filtered = false;
      for (int k = Lw + 1; k < L; ++k)
        filtered |= plan_->ops[k].kind == kFilterCompute && plan_->ops[k].u == w;
      if (filtered) continue;
      const OpDesc& cw = plan_->ops[Lw];
      uint32_t bu = 0;
      for (int i = 0; i < op.nb; ++i) bu |= 1u << op.bconn[i];
      const uint32_t need = compute_bconn_[w] | (1u << w);
      if ((need & bu) != need) continue;
      bool implied = true;
      for (int i = 0; i < cw.nc && implied; ++i) {
        const Cond c = cw.cond[i];
        if (c.op == kNonEqual) continue;
        bool ok = false;
        for (int j = 0; j < op.nc; ++j)
          if (op.cond[j].op == c.op && (op.cond[j].v == c.v || op.cond[j].v == w)) ok = true;
        implied = ok;
      }
      if (!implied) continue;
      reuse_w_[L] = w;
      keep_reuse_[L - 1] = true;
      for (int i = 0; i < op.nb; ++i)
        if (!(compute_bconn_[w] & (1u << op.bconn[i]))) rem_[L][nrem_[L]++] = op.bconn[i];
    }
  }

  // Same backward lists and the same effective conditions (non-equalities
  // implied by list membership dropped).
  bool SameIntersection(const OpDesc& a, const OpDesc& b) const {
    uint32_t ba = 0, bb = 0;
    for (int i = 0; i < a.nb; ++i) ba |= 1u << a.bconn[i];
    for (int i = 0; i < b.nb; ++i) bb |= 1u << b.bconn[i];
    if (ba != bb) return false;
    auto effective = [&](const OpDesc& o, uint32_t* lt, uint32_t* gt, uint32_t* ne) {
      *lt = *gt = *ne = 0;
      for (int i = 0; i < o.nc; ++i) {
        const uint32_t bit = 1u << o.cond[i].v;
        if (o.cond[i].op == kLess) *lt |= bit;
        if (o.cond[i].op == kLarger) *gt |= bit;
        if (o.cond[i].op == kNonEqual && (g_->self_loops || !(ba & bit))) *ne |= bit;
      }
    };
    uint32_t l1, g1, n1, l2, g2, n2;
    effective(a, &l1, &g1, &n1);
    effective(b, &l2, &g2, &n2);
    return l1 == l2 && g1 == g2 && n1 == n2;
  }

  uint64_t ReadU64(const uint64_t* dptr) {
    SME_HIP(hipMemcpyAsync(ctx_->pinned(), dptr, sizeof(uint64_t), hipMemcpyDeviceToHost, st_));
    SME_HIP(hipStreamSynchronize(st_));
    stats_.d2h_syncs++;
    return ctx_->pinned()[0];
  }

  uint64_t Budget(int L) const {
    const uint64_t avail = ctx_->RunAvailable();
    uint64_t per = static_cast<uint64_t>(0.9 * static_cast<double>(avail) /
                                         static_cast<double>(plan_->nops - L));
    if (ctx_->tuning.batch_bytes) per = std::min<uint64_t>(per, ctx_->tuning.batch_bytes);
    return std::max<uint64_t>(per, 4096);
  }


  void Scan(const uint64_t* in, uint64_t n, uint64_t* out) {
    Buf<uint64_t> tmp(pool_, ScanTempWords(n));
    Timed t(ctx_, st_, kSlotScan, 24.0 * n, n);
    LaunchScanU64(in, n, out, tmp.get(), st_);
  }
  void Scan32(const uint32_t* in, uint64_t n, uint64_t* out) {
    Buf<uint64_t> tmp(pool_, ScanTempWords(n));
    Timed t(ctx_, st_, kSlotScan, 20.0 * n, n);
    LaunchScanU32(in, n, out, tmp.get(), st_);
  }

  // Greedy split of paths [0, P) into batches whose cost
  //   items * bi64 / 64 + paths * bp
  // fits the level budget.  Returns right ends and the item offset at each end.
  void Split(const uint64_t* offs, uint64_t P, uint64_t W, uint64_t bi64, uint64_t bp, int L,
             std::vector<uint64_t>* ends, std::vector<uint64_t>* item_ends) {
    const uint64_t budget = Budget(L);
    const double total = static_cast<double>(W) * bi64 / 64.0 + static_cast<double>(P) * bp;
    ends->clear();
    item_ends->clear();
    if (total <= static_cast<double>(budget) && W <= kMaxBatchItems && P <= kMaxBatchPaths) {
      ends->push_back(P);
      item_ends->push_back(W);
      return;
    }
    const uint64_t want = static_cast<uint64_t>(total / budget) * 2 + 64;
    const uint64_t max_out = std::min<uint64_t>(want, P);
    // layout: [count, end_0, items_0, end_1, items_1, ...]
    Buf<uint64_t> bounds(pool_, 2 * max_out + 2);
    std::vector<uint64_t> host(2 * max_out);
    uint64_t start = 0;
    while (start < P) {
      {
        Timed t(ctx_, st_, kSlotMisc, 0, 0);
        LaunchSplitBatches(offs, start, P, bi64, bp * 64, budget * 64, kMaxBatchItems, kMaxBatchPaths,
                           bounds.get() + 1, max_out, bounds.get(), st_);
      }
      const uint64_t k = ReadU64(bounds.get());
      if (k == 0) Fail(kErrInvalid, "batch split made no progress");
      SME_HIP(hipMemcpyAsync(host.data(), bounds.get() + 1, 2 * k * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, st_));
      SME_HIP(hipStreamSynchronize(st_));
      stats_.d2h_syncs++;
      for (uint64_t i = 0; i < k; ++i) {
        ends->push_back(host[2 * i]);
        item_ends->push_back(host[2 * i + 1]);
      }
      start = ends->back();
    }
  }

  void Next(int L, const Frame& f) {
    if (L + 1 < plan_->nops && f.P > 0) Level(L + 1, f);
  }

  // Per-item temporaries of a verify/compact pass, in 1/64 byte: mask (8),
  // per-chunk count (4), per-chunk rank base (8), chunk partition (4), each
  // per 64 items.
  static constexpr uint64_t kChunkBytes64 = 8 + 4 + 8 + 4;
  // Per-path arrays the next level allocates up front (size, pstart, offs,
  // pivot id, scan scratch).
  uint64_t NextPathBytes(int L) const { return L + 1 < plan_->nops ? 32 : 0; }

  void Level0(uint32_t v_begin, uint32_t v_end) {
    const OpDesc& op = plan_->ops[0];
    const uint64_t total = v_end - v_begin;
    // the next level keeps ~48 bytes of per-path state
    const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(Budget(0) / 48, kMaxBatchPaths));
    for (uint64_t a = 0; a < total; a += per) {
      const uint64_t b = std::min(total, a + per);
      Frame f;
      f.P = b - a;
      f.cs.iota[op.u] = v_begin + static_cast<uint32_t>(a);
      f.matmask = 1u << op.u;
      host_level_[0] += f.P;
      stats_.batches++;
      Next(0, f);
    }
  }

  void Level(int L, const Frame& f) {
    const OpDesc& op = plan_->ops[L];
    switch (op.kind) {
      case kCompute:
      case kComputePathCount:
      case kComputeCount: ComputeLevel(L, f); break;
      case kMaterialize: MaterializeLevel(L, f); break;
      case kFilterCompute: FilterLevel(L, f); break;
      case kCount: CountLevel(L, f); break;
      default: Fail(kErrInvalid, "unknown op");
    }
  }

  static void WindowConds(const OpDesc& op, uint8_t* nlt, uint8_t* lt, uint8_t* ngt, uint8_t* gt) {
    *nlt = *ngt = 0;
    for (int i = 0; i < op.nc; ++i) {
      if (op.cond[i].op == kLess) lt[(*nlt)++] = op.cond[i].v;
      if (op.cond[i].op == kLarger) gt[(*ngt)++] = op.cond[i].v;
    }
  }

  // Non-equality conditions a candidate can actually violate.  x != M[w] is
  // implied when x was drawn from N(M[w]) and the graph has no self loops.
  uint8_t RealNonEqual(const OpDesc& op, uint32_t implied_mask, uint8_t* out) const {
    uint8_t n = 0;
    for (int i = 0; i < op.nc; ++i) {
      if (op.cond[i].op != kNonEqual) continue;
      const int w = op.cond[i].v;
      if (!g_->self_loops && (implied_mask & (1u << w))) continue;
      out[n++] = static_cast<uint8_t>(w);
    }
    return n;
  }

  ItemArgs BaseItems(const Frame& f, uint64_t l, uint64_t r, uint64_t items, const uint64_t* offs,
                     const uint32_t* chunk_seg, uint64_t nchunks) const {
    ItemArgs it{};
    it.nitems = items;
    it.nseg = r - l;
    it.nchunks = nchunks;
    it.offs = offs + l;
    it.chunk_seg = chunk_seg;
    it.rp = g_->rp;
    it.hub_id = g_->nhub ? g_->hub_id : nullptr;
    it.hub_bits = g_->hub_bits;
    it.hub_words = g_->hub_words;
    it.cols = g_->cols;
    it.cs = SliceCols(f.cs, l);
    it.super_chunks = ctx_->tuning.super_chunks;
    return it;
  }

  static void SetPairExclusion(const Frame& f, int w, uint64_t l, ItemArgs* it) {
    if (f.excl_u != w || !f.excl_in) return;
    it->pair_nexcl = f.nexcl;
    it->pair_excl_v[0] = f.excl_v[0];
    it->pair_excl_v[1] = f.excl_v[1];
    it->pair_excl_in = f.excl_in + l;
  }

  // A FILTER_COMPUTE whose output only feeds the plan's final pair count
  // (COMPUTE_COUNT with that one unmaterialized vertex) and whose real
  // conditions beyond the window are at most two != conditions: the count
  // can subtract the excluded values instead of reading a filtered copy.
  bool FilterAsView(int L, uint8_t nne) const {
    if (nne == 0 || nne > 2 || L + 2 != plan_->nops) return false;
    const OpDesc& nx = plan_->ops[L + 1];
    return nx.kind == kComputeCount && nx.ncun == 1 && nx.cun[0] == plan_->ops[L].u;
  }

  // --------------------------------------------------------------------------
  void ComputeLevel(int L, const Frame& f) {
    const OpDesc& op = plan_->ops[L];
    const uint64_t P = f.P;
    if (alias_of_[L] >= 0 && (f.candmask & (1u << plan_->ops[alias_of_[L]].u))) {
      // same set as an earlier level of this frame: share it
      Frame f2 = f;
      f2.has_reuse = false;
      f2.cand[op.u] = f.cand[plan_->ops[alias_of_[L]].u];
      f2.candmask |= 1u << op.u;
      CountPathArgs ca{};
      ca.npaths = P;
      ca.mode = 1;
      ca.ncun = 1;
      ca.cun[0] = op.u;
      ca.cand[op.u] = f2.cand[op.u];
      ca.acc_level = acc_ + 1 + L;
      {
        Timed t(ctx_, st_, kSlotCount, P * 12.0, P);
        LaunchCountPaths(ca, st_);
      }
      stats_.batches++;
      Next(L, f2);
      return;
    }
    // lists to intersect: the adjacency lists of `lists`, plus the parent's
    // candidate segment when this level extends it
    const bool reuse = reuse_w_[L] >= 0 && f.has_reuse;
    const uint8_t nb = reuse ? nrem_[L] : op.nb;
    const uint8_t* lists = reuse ? rem_[L] : op.bconn;
    const int nlists = nb + (reuse ? 1 : 0);
    Buf<uint64_t> size(pool_, P), pstart(pool_, P), offs(pool_, P + 1);
    Buf<uint8_t> piv(pool_, nlists > 1 ? P : 0);
    PrepComputeArgs pa{};
    pa.npaths = P;
    pa.rp = g_->rp;
    pa.cols = g_->cols;
    pa.cs = f.cs;
    pa.nb = nb;
    std::memcpy(pa.bconn, lists, kQ);
    if (reuse) {
      pa.has_reuse = 1;
      pa.reuse = f.reuse;
    }
    WindowConds(op, &pa.nlt, pa.lt_v, &pa.ngt, pa.gt_v);
    pa.size = size.get();
    pa.pstart = pstart.get();
    pa.piv = piv.get();
    pa.stream_ratio = ctx_->tuning.stream_ratio;

    uint32_t bmask = 0;
    for (int i = 0; i < op.nb; ++i) bmask |= 1u << op.bconn[i];
    uint8_t ne_v[kQ];
    const uint8_t nne = RealNonEqual(op, bmask, ne_v);
    const bool pair = op.kind == kComputeCount && op.ncun == 1;
    const int u = op.u;
    const bool zero_copy = nlists == 1 && nne == 0 && !pair;
    const double prep_bytes = P * (4.0 * (nlists + pa.nlt + pa.ngt) + 16.0 * nlists + 8.0 + 17.0);
    // paths extending a shared segment with one list: one wave per group
    const bool group = op.kind == kComputeCount && reuse && nb == 1 && !pair && nne <= 2 &&
                       ctx_->tuning.group_max > 0;
    if (group) {
      pa.group_cap = std::min<uint32_t>(ctx_->tuning.group_max, kGroupMax);
      GroupCountArgs ga{};
      ga.npaths = P;
      ga.cs = f.cs;
      ga.seg = f.reuse;
      ga.adj_v = lists[0];
      ga.rp = g_->rp;
      ga.cols = g_->cols;
      ga.up = g_->up;
      WindowConds(op, &ga.nlt, ga.lt_v, &ga.ngt, ga.gt_v);
      ga.nne = nne;
      ga.ne_v[0] = nne > 0 ? ne_v[0] : 0;
      ga.ne_v[1] = nne > 1 ? ne_v[1] : 0;
      ga.cap = pa.group_cap;
      ga.acc = acc_;
      ga.acc_level = acc_ + 1 + L;
      if (ctx_->profiling) {
        ga.walked = acc_ + 1 + 2 * plan_->nops + L;
        ga.lower = acc_ + 1 + 3 * plan_->nops + L;
      }
      ga.bk = bk_;
      // per path: gathered columns, two list bounds and the segment index;
      // the per-item bytes are added from the device counters after the run
      Timed t(ctx_, st_, kSlotGroup, P * (4.0 * (1 + ga.nlt + ga.ngt + nne) + 32.0 + 4.0), 0);
      LaunchGroupCount(ga, ctx_->tuning.item_grid, st_);
    }
    // long paths of a counting level: one workgroup per path
    const bool wg = op.kind == kComputeCount && !zero_copy && ctx_->tuning.wg_path_min > 0 &&
                    nlists - 1 <= 2 && nne <= 2;
    Buf<uint64_t> wg_list(pool_, wg ? 2 * P : 0);
    Buf<unsigned long long> wg_count(pool_, wg ? 1 : 0);
    if (wg) {
      LaunchFill64(reinterpret_cast<uint64_t*>(wg_count.get()), 1, 0, st_);
      pa.wg_min = ctx_->tuning.wg_path_min;
      pa.wg_list = wg_list.get();
      pa.wg_count = wg_count.get();
    }
    // short paths of a counting level: one lane per path in the set-up kernel
    const bool lane_count =
        op.kind == kComputeCount && !zero_copy && !group && ctx_->tuning.lane_path_max > 0;
    if (lane_count) {
      PrepCountArgs pc{};
      pc.pa = pa;
      pc.it = BaseItems(f, 0, P, 0, nullptr, nullptr, 0);
      pc.it.nb = nb;
      std::memcpy(pc.it.bconn, lists, kQ);
      if (reuse) {
        pc.it.has_reuse = 1;
        pc.it.reuse = f.reuse;
      }
      pc.it.nne = nne;
      std::memcpy(pc.it.ne_v, ne_v, sizeof(pc.it.ne_v));
      if (pair) {
        pc.it.pair = f.cand[op.cun[0]];
        pc.it.pair_op = op.pair_op;
        SetPairExclusion(f, op.cun[0], 0, &pc.it);
      }
      pc.it.acc = acc_;
      pc.it.acc_level = acc_ + 1 + L;
      pc.it.bk = bk_;
      pc.lane_max = ctx_->tuning.lane_path_max;
      pc.pair = pair ? 1 : 0;
      pc.lane_items = acc_ + 1 + plan_->nops + L;
      lane_lists_[L] = nlists;
      Timed t(ctx_, st_, kSlotLane, prep_bytes + (pair ? 20.0 * P : 0), P);
      LaunchPrepCount(pc, st_);
    } else {
      Timed t(ctx_, st_, kSlotPrep, prep_bytes, P);
      LaunchPrepCompute(pa, st_);
    }
    Scan(size.get(), P, offs.get());
    if (wg) {
      ItemArgs it = BaseItems(f, 0, P, 0, nullptr, nullptr, 0);
      it.src = g_->cols;
      it.pstart = pstart.get();
      it.piv = nlists > 1 ? piv.get() : nullptr;
      it.nb = nb;
      std::memcpy(it.bconn, lists, kQ);
      if (reuse) {
        it.has_reuse = 1;
        it.reuse = f.reuse;
      }
      it.nne = nne;
      std::memcpy(it.ne_v, ne_v, sizeof(it.ne_v));
      if (pair) {
        it.pair = f.cand[op.cun[0]];
        it.pair_op = op.pair_op;
        SetPairExclusion(f, op.cun[0], 0, &it);
      }
      it.acc = acc_;
      it.acc_level = acc_ + 1 + L;
      it.bk = bk_;
      wg_lists_[L] = nlists;
      Timed t(ctx_, st_, kSlotIntersect, 0, 0);
      LaunchPathWorkgroup(it, pair ? kModeSumPair : kModeSum, wg_list.get(), wg_count.get(),
                          acc_ + 1 + 4 * plan_->nops + L, ctx_->tuning.item_grid, st_);
    }
    const uint64_t W = ReadU64(offs.get() + P);
    if ((lane_count || wg || group) && W == 0) return;

    if (zero_copy) {
      // the candidate set is a window of one adjacency list: no copy
      if (op.kind == kComputeCount) {
        if (!bk_.bounds) {
          host_level_[L] += W;
          host_count_ += W;
          return;
        }
        // per-bucket: every path adds its window size to its bucket
        CountPathArgs ca{};
        ca.npaths = P;
        ca.mode = 1;
        ca.ncun = 1;
        ca.cun[0] = static_cast<uint8_t>(u);
        ca.cand[u] = CandRef{g_->cols, pstart.get(), size.get(), nullptr};
        ca.acc = acc_;
        ca.acc_level = acc_ + 1 + L;
        ca.cs = f.cs;
        ca.bk = bk_;
        Timed t(ctx_, st_, kSlotCount, P * 16.0, P);
        LaunchCountPaths(ca, st_);
        return;
      }
      Frame f2 = f;
      f2.has_reuse = false;
      f2.cand[u] = CandRef{g_->cols, pstart.get(), size.get(), nullptr};
      f2.candmask |= 1u << u;
      host_level_[L] += W;
      stats_.batches++;
      Next(L, f2);
      return;
    }

    // batch cost: per item (x1/64 byte) and per path, including the next
    // level's per-path set-up for the paths that continue
    const uint64_t next = NextPathBytes(L);
    uint64_t bi64 = 4, bp = 0;
    if (op.kind == kCompute) {
      bi64 = 64 * 4 + kChunkBytes64;
      bp = 16 + next;
    } else if (op.kind == kComputePathCount) {
      bi64 = kChunkBytes64;
      bp = 8 + next;
    }
    std::vector<uint64_t> ends, item_ends;
    Split(offs.get(), P, W, bi64, bp, L, &ends, &item_ends);
    uint64_t l = 0, il = 0;
    for (size_t b = 0; b < ends.size(); ++b) {
      const uint64_t r = ends[b], ir = item_ends[b];
      const uint64_t items = ir - il, nseg = r - l;
      const uint64_t nchunks = (items + 63) / 64;
      stats_.batches++;
      Buf<uint32_t> chunk_seg(pool_, nchunks);
      {
        Timed t(ctx_, st_, kSlotPartition, 16.0 * nseg + 4.0 * nchunks, nseg);
        LaunchChunkPartition(offs.get() + l, nseg, nchunks, chunk_seg.get(), st_);
      }
      ItemArgs it = BaseItems(f, l, r, items, offs.get(), chunk_seg.get(), nchunks);
      it.src = g_->cols;
      it.pstart = pstart.get() + l;
      it.piv = nlists > 1 ? piv.get() + l : nullptr;
      it.nb = nb;
      std::memcpy(it.bconn, lists, kQ);
      if (reuse) {
        it.has_reuse = 1;
        it.reuse = SliceCand(f.reuse, l);
      }
      it.nne = nne;
      std::memcpy(it.ne_v, ne_v, sizeof(it.ne_v));
      const double seg_bytes = 17.0 + 20.0 * (nlists - 1) + 4.0 * nne;
      if (op.kind == kComputeCount) {
        if (pair) {
          it.pair = SliceCand(f.cand[op.cun[0]], l);
          it.pair_op = op.pair_op;
          SetPairExclusion(f, op.cun[0], l, &it);
        }
        it.acc = acc_;
        it.acc_level = acc_ + 1 + L;
        it.bk = bk_;
        Timed t(ctx_, st_, kSlotIntersect, 4.0 * nlists * items + seg_bytes * nseg + (pair ? 20.0 * nseg : 0), items);
        LaunchVerify(it, pair ? kModeSumPair : kModeSum, ctx_->tuning.item_grid, st_);
        l = r;
        il = ir;
        continue;
      }
      Buf<unsigned long long> mask(pool_, nchunks);
      Buf<uint32_t> ccount(pool_, nchunks);
      it.mask = mask.get();
      it.ccount = ccount.get();
      {
        Timed t(ctx_, st_, kSlotIntersect, 4.0 * nlists * items + seg_bytes * nseg + 12.0 * nchunks, items);
        LaunchVerify(it, kModeMask, ctx_->tuning.item_grid, st_);
      }
      Buf<uint64_t> cbase(pool_, nchunks + 1);
      Scan32(ccount.get(), nchunks, cbase.get());
      const uint64_t C = ReadU64(cbase.get() + nchunks);
      RankRef rk{mask.get(), reinterpret_cast<const unsigned long long*>(cbase.get()), nchunks, items};
      const bool write_vals = op.kind == kCompute;
      Buf<uint32_t> vals(pool_, write_vals ? C : 0);
      Buf<uint64_t> beg(pool_, write_vals ? nseg : 0), len(pool_, nseg);
      CompactArgs ca{it, rk, vals.get(), beg.get(), len.get()};
      {
        Timed t(ctx_, st_, kSlotCompact, (write_vals ? 8.0 * C + 16.0 * nchunks : 0) + 32.0 * nseg, C);
        LaunchCompact(ca, st_);
      }
      // the rank structures are not needed below this level
      cbase.Reset();
      ccount.Reset();
      mask.Reset();
      chunk_seg.Reset();
      host_level_[L] += C;
      Frame f2 = SliceFrame(f, l, r);
      f2.has_reuse = false;
      f2.cand[u] = CandRef{vals.get(), beg.get(), len.get(), nullptr};
      f2.candmask |= 1u << u;
      Next(L, f2);
      l = r;
      il = ir;
    }
  }

  // --------------------------------------------------------------------------
  void MaterializeLevel(int L, const Frame& f) {
    const OpDesc& op = plan_->ops[L];
    const int u = op.u;
    const uint64_t P = f.P;
    if (!(f.candmask & (1u << u))) Fail(kErrInvalid, "MATERIALIZE of a vertex with no candidates");
    const CandRef cand = f.cand[u];
    Buf<uint64_t> size(pool_, P), pstart(pool_, P), offs(pool_, P + 1);
    PrepListArgs pa{};
    pa.npaths = P;
    pa.cs = f.cs;
    pa.cand = cand;
    WindowConds(op, &pa.nlt, pa.lt_v, &pa.ngt, pa.gt_v);
    pa.size = size.get();
    pa.pstart = pstart.get();
    {
      Timed t(ctx_, st_, kSlotPrep, P * (4.0 * (pa.nlt + pa.ngt) + 28.0), P);
      LaunchPrepList(pa, st_);
    }
    Scan(size.get(), P, offs.get());
    const uint64_t W = ReadU64(offs.get() + P);
    uint8_t ne_v[kQ];
    const uint8_t nne = RealNonEqual(op, compute_bconn_[u], ne_v);

    // columns the deeper levels still read
    uint8_t carry_v[kQ], ncarry = 0;
    for (int i = 0; i < op.nmat; ++i) {
      const int w = op.mat[i];
      if (live_after_[L] & (1u << w)) carry_v[ncarry++] = static_cast<uint8_t>(w);
    }
    uint8_t idx_v[kQ], nidx = 0;
    for (int w = 0; w < kQ; ++w)
      if (w != u && (f.candmask & (1u << w))) idx_v[nidx++] = static_cast<uint8_t>(w);

    // the next level extends u's candidate segment: children keep the index
    // of their parent's segment (one more index column)
    const bool keep = keep_reuse_[L];
    const uint8_t nidx_all = static_cast<uint8_t>(nidx + (keep ? 1 : 0));
    // every surviving item becomes a path of the next level
    const uint64_t bi64 = 64 * (4 * (1 + ncarry + nidx_all) + NextPathBytes(L)) + kChunkBytes64;
    std::vector<uint64_t> ends, item_ends;
    Split(offs.get(), P, W, bi64, 0, L, &ends, &item_ends);
    uint64_t l = 0, il = 0;
    for (size_t b = 0; b < ends.size(); ++b) {
      const uint64_t r = ends[b], ir = item_ends[b];
      const uint64_t items = ir - il, nseg = r - l;
      const uint64_t nchunks = (items + 63) / 64;
      stats_.batches++;
      Buf<uint32_t> chunk_seg(pool_, nchunks);
      {
        Timed t(ctx_, st_, kSlotPartition, 16.0 * nseg + 4.0 * nchunks, nseg);
        LaunchChunkPartition(offs.get() + l, nseg, nchunks, chunk_seg.get(), st_);
      }
      ItemArgs it = BaseItems(f, l, r, items, offs.get(), chunk_seg.get(), nchunks);
      it.src = cand.base;
      it.pstart = pstart.get() + l;
      it.nb = 0;
      it.nne = nne;
      std::memcpy(it.ne_v, ne_v, sizeof(it.ne_v));
      uint64_t C = items;
      Buf<unsigned long long> mask(pool_, nne ? nchunks : 0);
      Buf<uint32_t> ccount(pool_, nne ? nchunks : 0);
      Buf<uint64_t> cbase(pool_, nne ? nchunks + 1 : 0);
      RankRef rk{nullptr, nullptr, nchunks, items};
      if (nne) {
        it.mask = mask.get();
        it.ccount = ccount.get();
        {
          Timed t(ctx_, st_, kSlotFilter, 4.0 * items + (16.0 + 4.0 * nne) * nseg + 12.0 * nchunks, items);
          LaunchVerify(it, kModeMask, ctx_->tuning.item_grid, st_);
        }
        Scan32(ccount.get(), nchunks, cbase.get());
        C = ReadU64(cbase.get() + nchunks);
        rk.mask = mask.get();
        rk.cbase = reinterpret_cast<const unsigned long long*>(cbase.get());
      }
      Buf<uint32_t> out_new(pool_, C);
      std::vector<std::unique_ptr<Buf<uint32_t>>> carry, idx;
      ExpandArgs ea{};
      ea.it = it;
      ea.rk = rk;
      ea.out_new = out_new.get();
      ea.ncarry = ncarry;
      for (int i = 0; i < ncarry; ++i) {
        carry.emplace_back(new Buf<uint32_t>(pool_, C));
        ea.carry_v[i] = carry_v[i];
        ea.carry_out[i] = carry.back()->get();
      }
      ea.nidx = nidx_all;
      for (int i = 0; i < nidx_all; ++i) {
        idx.emplace_back(new Buf<uint32_t>(pool_, C));
        ea.idx_in[i] = SliceCand(i < nidx ? f.cand[idx_v[i]] : cand, l);
        ea.idx_out[i] = idx.back()->get();
      }
      {
        Timed t(ctx_, st_, kSlotExpand,
                4.0 * C * (2 + ncarry + nidx) + nseg * (16.0 + 4.0 * ncarry + 4.0 * nidx), C);
        LaunchExpand(ea, st_);
      }
      cbase.Reset();
      ccount.Reset();
      mask.Reset();
      chunk_seg.Reset();
      host_level_[L] += C;
      Frame f2;
      f2.P = C;
      f2.cs.inst[u] = out_new.get();
      f2.matmask = 1u << u;
      for (int i = 0; i < ncarry; ++i) {
        f2.cs.inst[carry_v[i]] = carry[i]->get();
        f2.matmask |= 1u << carry_v[i];
      }
      for (int i = 0; i < nidx; ++i) {
        CandRef c = ea.idx_in[i];
        c.idx = idx[i]->get();
        f2.cand[idx_v[i]] = c;
        f2.candmask |= 1u << idx_v[i];
      }
      if (keep) {
        f2.reuse = ea.idx_in[nidx];
        f2.reuse.idx = idx[nidx]->get();
        f2.has_reuse = true;
      }
      const bool gpsm = plan_->algo == kAlgoGpsm || plan_->algo == kAlgoGpsmReuse;
      if (gpsm && list_ && L + 1 == plan_->nops && C && list_rows_ < list_cap_) {
        const uint64_t rows = std::min<uint64_t>(C, list_cap_ - list_rows_);
        Buf<uint32_t> tuples(pool_, rows * plan_->nv);
        {
          Timed t(ctx_, st_, kSlotExpand, 4.0 * rows * plan_->nv * 2, rows);
          LaunchGatherTuples(f2.cs, plan_->nv, rows, tuples.get(), st_);
        }
        SME_HIP(hipMemcpyAsync(list_out_ + list_rows_ * plan_->nv, tuples.get(),
                               rows * plan_->nv * sizeof(uint32_t), hipMemcpyDeviceToHost, st_));
        SME_HIP(hipStreamSynchronize(st_));
        stats_.d2h_syncs++;
        list_rows_ += rows;
      }
      if (gpsm && bk_.bounds && L + 1 == plan_->nops && C) {
        // GPSM counts the last level's partial instances: one per path
        CountPathArgs ca{};
        ca.npaths = C;
        ca.mode = 2;
        ca.cs = f2.cs;
        ca.bk = bk_;
        Timed t(ctx_, st_, kSlotCount, C * 8.0, C);
        LaunchCountPaths(ca, st_);
      }
      Next(L, f2);
      l = r;
      il = ir;
    }
  }

  // --------------------------------------------------------------------------
  void FilterLevel(int L, const Frame& f) {
    const OpDesc& op = plan_->ops[L];
    const int u = op.u;
    const uint64_t P = f.P;
    if (!(f.candmask & (1u << u))) Fail(kErrInvalid, "FILTER_COMPUTE of a vertex with no candidates");
    const CandRef cand = f.cand[u];
    Buf<uint64_t> size(pool_, P), pstart(pool_, P), offs(pool_, P + 1);
    PrepListArgs pa{};
    pa.npaths = P;
    pa.cs = f.cs;
    pa.cand = cand;
    WindowConds(op, &pa.nlt, pa.lt_v, &pa.ngt, pa.gt_v);
    pa.size = size.get();
    pa.pstart = pstart.get();
    {
      Timed t(ctx_, st_, kSlotPrep, P * (4.0 * (pa.nlt + pa.ngt) + 28.0), P);
      LaunchPrepList(pa, st_);
    }
    Scan(size.get(), P, offs.get());
    const uint64_t W = ReadU64(offs.get() + P);
    uint8_t ne_v[kQ];
    const uint8_t nne = RealNonEqual(op, compute_bconn_[u], ne_v);
    if (nne == 0) {
      Frame f2 = f;
      f2.cand[u] = CandRef{cand.base, pstart.get(), size.get(), nullptr};
      host_level_[L] += W;
      stats_.batches++;
      Next(L, f2);
      return;
    }
    if (FilterAsView(L, nne)) {
      Buf<uint8_t> excl(pool_, P);
      ExclusionArgs ea{};
      ea.npaths = P;
      ea.cs = f.cs;
      ea.cand = CandRef{cand.base, pstart.get(), size.get(), nullptr};
      ea.nne = nne;
      ea.ne_v[0] = ne_v[0];
      ea.ne_v[1] = nne > 1 ? ne_v[1] : 0;
      ea.excl_in = excl.get();
      ea.acc_level = acc_ + 1 + L;
      {
        Timed t(ctx_, st_, kSlotFilter, P * (17.0 + 4.0 * nne), P);
        LaunchExclusion(ea, st_);
      }
      Frame f2 = f;
      f2.cand[u] = ea.cand;
      f2.excl_u = u;
      f2.nexcl = nne;
      f2.excl_v[0] = ea.ne_v[0];
      f2.excl_v[1] = ea.ne_v[1];
      f2.excl_in = excl.get();
      stats_.batches++;
      Next(L, f2);
      return;
    }
    std::vector<uint64_t> ends, item_ends;
    Split(offs.get(), P, W, 64 * 4 + kChunkBytes64, 16 + NextPathBytes(L), L, &ends, &item_ends);
    uint64_t l = 0, il = 0;
    for (size_t b = 0; b < ends.size(); ++b) {
      const uint64_t r = ends[b], ir = item_ends[b];
      const uint64_t items = ir - il, nseg = r - l;
      const uint64_t nchunks = (items + 63) / 64;
      stats_.batches++;
      Buf<uint32_t> chunk_seg(pool_, nchunks);
      {
        Timed t(ctx_, st_, kSlotPartition, 16.0 * nseg + 4.0 * nchunks, nseg);
        LaunchChunkPartition(offs.get() + l, nseg, nchunks, chunk_seg.get(), st_);
      }
      ItemArgs it = BaseItems(f, l, r, items, offs.get(), chunk_seg.get(), nchunks);
      it.src = cand.base;
      it.pstart = pstart.get() + l;
      it.nb = 0;
      it.nne = nne;
      std::memcpy(it.ne_v, ne_v, sizeof(it.ne_v));
      Buf<unsigned long long> mask(pool_, nchunks);
      Buf<uint32_t> ccount(pool_, nchunks);
      it.mask = mask.get();
      it.ccount = ccount.get();
      {
        Timed t(ctx_, st_, kSlotFilter, 4.0 * items + (16.0 + 4.0 * nne) * nseg + 12.0 * nchunks, items);
        LaunchVerify(it, kModeMask, ctx_->tuning.item_grid, st_);
      }
      Buf<uint64_t> cbase(pool_, nchunks + 1);
      Scan32(ccount.get(), nchunks, cbase.get());
      const uint64_t C = ReadU64(cbase.get() + nchunks);
      RankRef rk{mask.get(), reinterpret_cast<const unsigned long long*>(cbase.get()), nchunks, items};
      Buf<uint32_t> vals(pool_, C);
      Buf<uint64_t> beg(pool_, nseg), len(pool_, nseg);
      CompactArgs ca{it, rk, vals.get(), beg.get(), len.get()};
      {
        Timed t(ctx_, st_, kSlotCompact, 8.0 * C + 16.0 * nchunks + 32.0 * nseg, C);
        LaunchCompact(ca, st_);
      }
      cbase.Reset();
      ccount.Reset();
      mask.Reset();
      chunk_seg.Reset();
      host_level_[L] += C;
      Frame f2 = SliceFrame(f, l, r);
      f2.cand[u] = CandRef{vals.get(), beg.get(), len.get(), nullptr};
      Next(L, f2);
      l = r;
      il = ir;
    }
  }

  // --------------------------------------------------------------------------
  void CountLevel(int L, const Frame& f) {
    const OpDesc& op = plan_->ops[L];
    const uint64_t P = f.P;
    if (plan_->direct_count || op.ncun == 1) {
      CountPathArgs ca{};
      ca.npaths = P;
      ca.mode = plan_->direct_count ? 0 : 1;
      ca.query_type = plan_->query_type;
      ca.ncun = op.ncun;
      std::memcpy(ca.cun, op.cun, sizeof(ca.cun));
      for (int i = 0; i < op.ncun; ++i) ca.cand[op.cun[i]] = f.cand[op.cun[i]];
      ca.acc = acc_;
      ca.acc_level = acc_ + 1 + L;
      ca.cs = f.cs;
      ca.bk = bk_;
      Timed t(ctx_, st_, kSlotCount, P * 20.0 * op.ncun, P);
      LaunchCountPaths(ca, st_);
      return;
    }
    if (op.ncun != 2) Fail(kErrUnsupported, "COUNT needs one or two unmaterialized vertices");
    const CandRef a = f.cand[op.cun[0]], b = f.cand[op.cun[1]];
    Buf<uint64_t> size(pool_, P), pstart(pool_, P), offs(pool_, P + 1);
    Buf<uint8_t> swapped(pool_, P);
    PrepTwoArgs pa{P, a, b, size.get(), pstart.get(), swapped.get()};
    {
      Timed t(ctx_, st_, kSlotPrep, P * 57.0, P);
      LaunchPrepTwo(pa, st_);
    }
    Scan(size.get(), P, offs.get());
    const uint64_t W = ReadU64(offs.get() + P);
    std::vector<uint64_t> ends, item_ends;
    Split(offs.get(), P, W, 4, 0, L, &ends, &item_ends);
    uint64_t l = 0, il = 0;
    for (size_t bt = 0; bt < ends.size(); ++bt) {
      const uint64_t r = ends[bt], ir = item_ends[bt];
      const uint64_t items = ir - il, nseg = r - l;
      const uint64_t nchunks = (items + 63) / 64;
      stats_.batches++;
      Buf<uint32_t> chunk_seg(pool_, nchunks);
      {
        Timed t(ctx_, st_, kSlotPartition, 16.0 * nseg + 4.0 * nchunks, nseg);
        LaunchChunkPartition(offs.get() + l, nseg, nchunks, chunk_seg.get(), st_);
      }
      ItemArgs it = BaseItems(f, l, r, items, offs.get(), chunk_seg.get(), nchunks);
      it.pstart = pstart.get() + l;
      it.pair = SliceCand(a, l);
      it.second = SliceCand(b, l);
      it.swapped = swapped.get() + l;
      it.pair_op = op.pair_op;
      it.acc = acc_;
      it.acc_level = acc_ + 1 + L;
      it.bk = bk_;
      Timed t(ctx_, st_, kSlotCount, 8.0 * items + 40.0 * nseg, items);
      LaunchVerify(it, kModeSumTwo, ctx_->tuning.item_grid, st_);
      l = r;
      il = ir;
    }
  }

  Context* ctx_;
  const DevGraph* g_;
  const EnginePlan* plan_;
  hipStream_t st_;
  Context* pool_;  // allocations of this run go through the context
  unsigned long long* acc_ = nullptr;
  bool list_ = false;
  uint32_t* list_out_ = nullptr;
  uint64_t list_cap_ = 0, list_rows_ = 0;
  int lane_lists_[kMaxOps];  // lists per item of lane-counted levels
  int wg_lists_[kMaxOps];    // lists per item of workgroup-counted levels
  uint64_t host_level_[kMaxOps];
  uint64_t levels_[kMaxOps];
  uint64_t host_count_ = 0;
  uint32_t live_after_[kMaxOps];
  uint32_t compute_bconn_[kQ];
  BucketRef bk_{};
  int alias_of_[kMaxOps];        // earlier level computing the same set, or -1
  int reuse_w_[kMaxOps];         // vertex whose candidate set level L extends
  bool keep_reuse_[kMaxOps];     // MATERIALIZE at L keeps the parent segment
  uint8_t rem_[kMaxOps][kQ];     // backward lists the reused set does not cover
  uint8_t nrem_[kMaxOps];
  RunStats stats_;
};

// After a failed run: drain the stream so that pool memory handed back is no
// longer in use by queued kernels, and recycle the run's timing events.
Status Recover(Context* ctx, Status s, const std::string& msg, std::string* err) {
  (void)hipStreamSynchronize(ctx->stream());
  for (auto& m : ctx->marks) {
    ctx->spare_events.push_back(m.start);
    ctx->spare_events.push_back(m.stop);
  }
  ctx->marks.clear();
  *err = msg;
  return s;
}

// ---------------------------------------------------------------------------
// Clique plans over bitmaps of upper neighbourhoods (clique.hip).  Level L of
// a LIGHT clique plan with ordering outputs the j-cliques (j = vertices
// matched after L: acc[nmat + 1], whatever the op) whose lowest vertex is a
// start vertex, and its count is acc[k].
bool LocalCliqueApplies(const Context* ctx, const EnginePlan* plan) {
  return !ctx->tuning.clique_level_engine && plan->algo == kAlgoLight && plan->clique &&
         plan->enable_ordering && plan->nv >= 3 && plan->nv <= kCliqueMaxK;
}

void LocalCliqueRange(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t vb, uint32_t ve,
                      const BucketRef& bk, uint64_t* count, uint64_t* levels, RunStats* st,
                      uint32_t* tc = nullptr) {
  hipStream_t s = ctx->stream();
  const int K = plan->nv;
  const uint64_t range = ve > vb ? ve - vb : 0;
  constexpr int kAcc = kCliqueMaxK + 1;
  Buf<unsigned long long> acc(ctx, kAcc + 2);
  LaunchFill64(reinterpret_cast<uint64_t*>(acc.get()), kAcc + 2, 0, s);
  // over all start vertices the upload-time lists serve; otherwise the small
  // kernel appends the range's larger vertices and they are sorted here
  const bool whole = vb == 0 && ve == g->n;
  Buf<uint32_t> big(ctx, whole ? 0 : range), big_d(ctx, whole ? 0 : range), huge(ctx, whole ? 0 : range);
  CliqueArgs a{};
  a.rp = g->rp;
  a.cols = g->cols;
  a.up = g->up;
  a.k = K;
  a.acc = acc.get();
  a.bk = bk;
  a.v_begin = vb;
  a.v_end = ve;
  a.big = big.get();
  a.big_d = big_d.get();
  a.nbig = acc.get() + kAcc;
  a.huge = huge.get();
  a.nhuge = acc.get() + kAcc + 1;
  a.tc = tc;
  {
    Timed t(ctx, s, kSlotCliqueSmall, 20.0 * static_cast<double>(range), range);
    LaunchCliqueSmall(a, s);
  }
  uint64_t* pin = ctx->pinned();
  st->batches = 1;
  uint64_t nbig = 0, nhuge = 0;
  if (whole) {
    nbig = g->upper.verts.size();
    nhuge = g->upper.huge.size();
  } else {
    SME_HIP(hipMemcpyAsync(pin, acc.get() + kAcc, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    SME_HIP(hipStreamSynchronize(s));
    st->d2h_syncs++;
    nbig = pin[0];
    nhuge = pin[1];
  }
  std::vector<uint32_t> hv, hd, verts_local, hh, hd_local;
  std::vector<uint64_t> rf_local, wf_local, cf_local;
  const std::vector<uint32_t>& verts = whole ? g->upper.verts : verts_local;
  const std::vector<uint32_t>& hd_sorted = whole ? g->upper.d : hd_local;
  const std::vector<uint64_t>& rf = whole ? g->upper.rf : rf_local;
  const std::vector<uint64_t>& wf = whole ? g->upper.wf : wf_local;
  const std::vector<uint64_t>& cf = whole ? g->upper.cf : cf_local;
  Buf<uint32_t> dverts_local(ctx, whole ? 0 : nbig);
  Buf<uint64_t> drf_local(ctx, whole || !nbig ? 0 : nbig + 1), dwf_local(ctx, whole || !nbig ? 0 : nbig + 1),
      dcf_local(ctx, whole || !nbig ? 0 : nbig + 1);
  const uint32_t* dverts = whole ? g->upper.dverts : dverts_local.get();
  const uint64_t* drf = whole ? g->upper.drf : drf_local.get();
  const uint64_t* dwf = whole ? g->upper.dwf : dwf_local.get();
  const uint64_t* dcf = whole ? g->upper.dcf : dcf_local.get();
  if (nbig && !whole) {
    hv.resize(nbig);
    hd.resize(nbig);
    SME_HIP(hipMemcpyAsync(hv.data(), big.get(), nbig * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SME_HIP(hipMemcpyAsync(hd.data(), big_d.get(), nbig * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SME_HIP(hipStreamSynchronize(s));
    st->d2h_syncs++;
    // largest neighbourhoods first (the append order is arbitrary)
    std::vector<uint32_t> order(nbig);
    for (uint64_t x = 0; x < nbig; ++x) order[x] = static_cast<uint32_t>(x);
    std::sort(order.begin(), order.end(), [&](uint32_t p, uint32_t q) {
      return hd[p] != hd[q] ? hd[p] > hd[q] : hv[p] < hv[q];
    });
    hd_local.resize(nbig);
    verts_local.resize(nbig);
    rf_local.assign(nbig + 1, 0);
    wf_local.assign(nbig + 1, 0);
    cf_local.assign(nbig + 1, 0);
    for (uint64_t x = 0; x < nbig; ++x) {
      const uint64_t d = hd[order[x]];
      hd_local[x] = hd[order[x]];
      verts_local[x] = hv[order[x]];
      rf_local[x + 1] = rf_local[x] + d;
      wf_local[x + 1] = wf_local[x] + d * ((d + 63) / 64);
      cf_local[x + 1] = cf_local[x] + (d + 63) / 64;
    }
    SME_HIP(hipMemcpyAsync(dcf_local.get(), cf.data(), (nbig + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    SME_HIP(hipMemcpyAsync(dverts_local.get(), verts.data(), nbig * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    SME_HIP(hipMemcpyAsync(drf_local.get(), rf.data(), (nbig + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    SME_HIP(hipMemcpyAsync(dwf_local.get(), wf.data(), (nbig + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  }
  if (nbig) {
    // batches of whole vertices whose rows fit what the run may use
    uint64_t cap_words = static_cast<uint64_t>(0.9 * static_cast<double>(ctx->RunAvailable())) / 8;
    if (ctx->tuning.batch_bytes) cap_words = std::min<uint64_t>(cap_words, ctx->tuning.batch_bytes / 8);
    st->batches = 0;
    // the largest neighbourhoods come first: any whose rows alone exceed
    // the cap join the level engine's start vertices below
    uint64_t first_fit = 0;
    while (first_fit < nbig && wf[first_fit + 1] - wf[first_fit] > cap_words) ++first_fit;
    if (first_fit) {
      if (tc) Fail(kErrOom, "bit rows of one start vertex exceed the memory the run may use");
      hh.insert(hh.end(), verts.begin(), verts.begin() + static_cast<int64_t>(first_fit));
    }
    for (uint64_t x0 = first_fit; x0 < nbig;) {
      uint64_t x1 = x0 + 1;
      while (x1 < nbig && wf[x1 + 1] - wf[x0] <= cap_words) ++x1;
      const uint64_t words = wf[x1] - wf[x0];
      Buf<unsigned long long> rows(ctx, words);
      a.verts = dverts + x0;
      a.nverts = x1 - x0;
      a.nitems = rf[x1] - rf[x0];
      a.row_first = drf + x0;
      a.word_first = dwf + x0;
      a.rows = rows.get();
      // the batch is in decreasing d: vertices above kRowsHashMax search,
      // the others stream against an LDS hash (a small table up to
      // kRowsHashSmall)
      const uint64_t hash_from = std::partition_point(hd_sorted.begin() + x0, hd_sorted.begin() + x1,
                                                      [](uint32_t d) { return d > kRowsHashMax; }) -
                                 hd_sorted.begin();
      const uint64_t small_from = std::partition_point(hd_sorted.begin() + hash_from, hd_sorted.begin() + x1,
                                                       [](uint32_t d) { return d > kRowsHashSmall; }) -
                                  hd_sorted.begin();
      {
        // per row: the row vertex, its list bounds (20 bytes) and its words
        // written; per later neighbour of the start vertex: the key and one
        // probe of the row vertex's list (4 + 4 bytes)
        double keys = 0;
        for (uint64_t x = x0; x < x1; ++x) {
          const double d = static_cast<double>(rf[x + 1] - rf[x]);
          keys += d * (d - 1) / 2;
        }
        Timed t(ctx, s, kSlotCliqueRows,
                8.0 * keys + 20.0 * static_cast<double>(a.nitems) + 8.0 * static_cast<double>(words), a.nitems);
        if (hash_from > x0) {
          CliqueArgs b = a;
          b.nverts = hash_from - x0;
          b.nitems = rf[hash_from] - rf[x0];
          LaunchCliqueRows(b, s);
        }
        for (int cls = 0; cls < 2; ++cls) {
          const uint64_t xa = cls ? small_from : hash_from, xb = cls ? x1 : small_from;
          if (xb <= xa) continue;
          CliqueArgs b = a;
          b.verts = dverts + xa;
          b.nverts = xb - xa;
          b.row_first = drf + xa;
          b.word_first = dwf + xa;
          b.chunk_first = dcf + xa;
          b.nchunks = cf[xb] - cf[xa];
          b.rows = rows.get() + (wf[xa] - wf[x0]);
          LaunchCliqueRowsHash(b, cls == 1, s);
        }
      }
      if (!tc) {
        // every row word read at least once
        Timed t(ctx, s, kSlotCliqueCount, 8.0 * static_cast<double>(words), a.nitems);
        LaunchCliqueCount(a, s);
      }
      st->batches++;
      x0 = x1;
    }
    SME_HIP(hipStreamSynchronize(s));
  }
  SME_HIP(hipMemcpyAsync(pin, acc.get(), kAcc * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  const uint64_t nfirst = hh.size();  // start vertices too large to batch
  hh.resize(nfirst + nhuge);
  if (nhuge && whole)
    std::copy(g->upper.huge.begin(), g->upper.huge.end(), hh.begin() + static_cast<int64_t>(nfirst));
  else if (nhuge)
    SME_HIP(hipMemcpyAsync(hh.data() + nfirst, huge.get(), nhuge * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  SME_HIP(hipStreamSynchronize(s));
  st->d2h_syncs++;
  uint64_t cl[kAcc];
  std::memcpy(cl, pin, sizeof(cl));
  // the level engine below counts those vertices whole (their 1- and
  // 2-cliques included)
  for (uint64_t x = 0; x < nfirst; ++x) {
    cl[1] -= 1;
    cl[2] -= hd_sorted[x];
  }
  uint64_t total = cl[K];
  if (levels)
    for (int L = 0; L < plan->nops; ++L) levels[L] = cl[std::min<int>(plan->ops[L].nmat + 1, K)];
  // neighbourhoods beyond kLocalMax: the level engine, one run per range of
  // consecutive such start vertices (per-edge triangle counts have no such
  // fallback: their callers check the graph's largest upper neighbourhood)
  if (tc && nhuge) Fail(kErrUnsupported, "upper neighbourhood beyond the bitmap path's limit");
  std::sort(hh.begin(), hh.end());
  for (uint64_t x = 0; x < hh.size();) {
    uint64_t y = x + 1;
    while (y < hh.size() && hh[y] == hh[y - 1] + 1) ++y;
    Runner r(ctx, g, plan);
    if (bk.bounds) r.SetBuckets(bk.bounds, bk.nbuckets, bk.counts, bk.table, bk.base);
    uint64_t c = 0, lv[kMaxOps] = {0};
    RunStats rs;
    r.Run(hh[x], hh[y - 1] + 1, &c, levels ? lv : nullptr, &rs);
    total += c;
    if (levels)
      for (int L = 0; L < plan->nops; ++L) levels[L] += lv[L];
    st->batches += rs.batches;
    st->d2h_syncs += rs.d2h_syncs;
    x = y;
  }
  *count = total;
}

// Start ranges whose vertex lists (12 bytes per start vertex) would take more
// than a quarter of what the run may use are counted in consecutive parts.
void LocalClique(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t vb, uint32_t ve,
                 const BucketRef& bk, uint64_t* count, uint64_t* levels, RunStats* st,
                 uint32_t* tc = nullptr) {
  const uint64_t part = std::max<uint64_t>(64, ctx->RunAvailable() / 4 / 12);
  *count = 0;
  if (levels) std::fill(levels, levels + plan->nops, 0);
  *st = RunStats{};
  uint64_t b = vb;
  do {
    const uint32_t e = static_cast<uint32_t>(std::min<uint64_t>(ve, b + part));
    uint64_t c = 0, lv[kMaxOps] = {0};
    RunStats rs;
    LocalCliqueRange(ctx, g, plan, static_cast<uint32_t>(b), e, bk, &c, levels ? lv : nullptr, &rs, tc);
    *count += c;
    if (levels)
      for (int L = 0; L < plan->nops; ++L) levels[L] += lv[L];
    st->batches += rs.batches;
    st->d2h_syncs += rs.d2h_syncs;
    b = e;
  } while (b < ve);
}

// ---------------------------------------------------------------------------
// House count over all start vertices from edge statistics (house.hip).  The
// packed accumulators hold a wedge count (at most the largest degree) above
// bit kHouseShift and a sum of at most that many pairs of triangle counts
// (each below the largest degree) below it, so the largest degree must stay
// under 2^19; the per-edge triangle pass needs every upper neighbourhood
// within the clique bitmap path's limit.
// Device bytes the per-edge triangle pass needs at least: the counts, the
// start-vertex lists of one part and the bit rows of the largest upper
// neighbourhood.
uint64_t CensusBytes(const DevGraph* g) {
  const uint64_t d = g->max_up_degree;
  return 4 * g->nnz + 12 * 4096 + d * ((d + 63) / 64) * 8 + (1 << 20);
}

bool LocalHouseApplies(const Context* ctx, const DevGraph* g, const EnginePlan* plan) {
  // plus the wedge loads and one dense accumulator with its touched list
  return !ctx->tuning.census_level_engine && plan->query_type == 4 && plan->enable_ordering &&
         !g->self_loops && g->max_degree < (1ull << 19) && g->max_up_degree <= kLocalMax &&
         CensusBytes(g) + 20 * g->n <= ctx->RunAvailable();
}

void LocalHouse(Context* ctx, const DevGraph* g, uint64_t* count, RunStats* st) {
  hipStream_t s = ctx->stream();
  const uint64_t n = g->n, nnz = g->nnz;
  *count = 0;
  *st = RunStats{};
  st->batches = 1;
  if (n == 0 || nnz == 0) return;
  // 1. triangles through every edge, recorded at its upper entry
  Buf<uint32_t> t(ctx, nnz);
  LaunchFill32(t.get(), nnz, 0, s);
  {
    EnginePlan tri;
    std::string ignored;
    if (!MakeLightEnginePlan(0, kClSpecial, true, &tri, nullptr)) Fail(kErrInvalid, "triangle plan");
    uint64_t c = 0;
    RunStats rs;
    LocalClique(ctx, g, &tri, 0, static_cast<uint32_t>(n), BucketRef{}, &c, nullptr, &rs, t.get());
    st->d2h_syncs += rs.d2h_syncs;
  }
  // 2. mirrored counts, wedge loads, sum of t(t-1)
  Buf<unsigned long long> acc(ctx, 2);
  LaunchFill64(reinterpret_cast<uint64_t*>(acc.get()), 2, 0, s);
  Buf<unsigned long long> load(ctx, n);
  HouseArgs h{};
  h.rp = g->rp;
  h.cols = g->cols;
  h.up = g->up;
  h.n = n;
  h.t = t.get();
  h.load = load.get();
  h.acc = acc.get();
  {
    Timed tm(ctx, s, kSlotHouseMirror, 12.0 * static_cast<double>(nnz) + 32.0 * static_cast<double>(n), nnz);
    LaunchHouseMirror(h, s);
  }
  std::vector<unsigned long long> hl(n);
  std::vector<uint64_t> hrp(n + 1), hup(n);
  SME_HIP(hipMemcpyAsync(hl.data(), load.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SME_HIP(hipMemcpyAsync(hrp.data(), g->rp, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SME_HIP(hipMemcpyAsync(hup.data(), g->up, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SME_HIP(hipStreamSynchronize(s));
  st->d2h_syncs++;
  std::vector<uint32_t> light, heavy;
  double wedges = 0;
  for (uint64_t u = 0; u < n; ++u) {
    if (!hl[u]) continue;
    wedges += static_cast<double>(hl[u]);
    (hl[u] <= kHouseHashMax ? light : heavy).push_back(static_cast<uint32_t>(u));
  }
  std::sort(heavy.begin(), heavy.end(), [&](uint32_t a, uint32_t b) {
    return hl[a] != hl[b] ? hl[a] > hl[b] : a < b;
  });
  // 3a. light apexes: one wave each, LDS hash
  if (!light.empty()) {
    Buf<uint32_t> d(ctx, light.size());
    SME_HIP(hipMemcpyAsync(d.get(), light.data(), light.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    h.apex = d.get();
    h.napex = light.size();
    double lw = 0;
    for (uint32_t u : light) lw += static_cast<double>(hl[u]);
    {
      // per wedge: the entry w and its t
      Timed tm(ctx, s, kSlotHouseWedge, 8.0 * lw, static_cast<uint64_t>(lw));
      LaunchHouseWedgeHash(h, s);
    }
    SME_HIP(hipStreamSynchronize(s));
  }
  // 3b. heavy apexes: groups of dense per-apex arrays over w
  if (!heavy.empty()) {
    const uint64_t per_slot = 12 * n;  // u64 accumulator + u32 touched entry per w
    const uint64_t fit = static_cast<uint64_t>(0.5 * static_cast<double>(ctx->RunAvailable())) / per_slot;
    const uint64_t S = std::max<uint64_t>(1, std::min<uint64_t>({256, heavy.size(), fit}));
    Buf<unsigned long long> slots(ctx, S * n);
    Buf<uint32_t> touched(ctx, S * n);
    Buf<unsigned long long> ntouched(ctx, S);
    Buf<uint32_t> dapex(ctx, S);
    Buf<uint64_t> dfirst(ctx, S + 1);
    LaunchFill64(reinterpret_cast<uint64_t*>(slots.get()), S * n, 0, s);
    LaunchFill64(reinterpret_cast<uint64_t*>(ntouched.get()), S, 0, s);
    std::vector<uint64_t> first(S + 1);
    for (uint64_t g0 = 0; g0 < heavy.size(); g0 += S) {
      const uint64_t g1 = std::min<uint64_t>(heavy.size(), g0 + S);
      const uint64_t na = g1 - g0;
      uint64_t nfull = 0;
      first[0] = 0;
      double gw = 0;
      for (uint64_t x = 0; x < na; ++x) {
        const uint32_t u = heavy[g0 + x];
        first[x + 1] = first[x] + (hup[u] - hrp[u]);
        // an apex touching a large share of [0, u) is reduced by a scan
        if (hl[u] > u / 8) nfull = x + 1;
        gw += static_cast<double>(hl[u]);
      }
      SME_HIP(hipMemcpyAsync(dapex.get(), heavy.data() + g0, na * sizeof(uint32_t), hipMemcpyHostToDevice, s));
      SME_HIP(hipMemcpyAsync(dfirst.get(), first.data(), (na + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
      h.apex = dapex.get();
      h.napex = na;
      h.item_first = dfirst.get();
      h.nitems = first[na];
      h.nfull = nfull;
      h.slots = slots.get();
      h.touched = touched.get();
      h.ntouched = ntouched.get();
      {
        Timed tm(ctx, s, kSlotHouseWedge, 8.0 * gw, static_cast<uint64_t>(gw));
        LaunchHouseWedgeDense(h, s);
        LaunchHouseWedgeReduce(h, s);
      }
      LaunchFill64(reinterpret_cast<uint64_t*>(ntouched.get()), na, 0, s);
      SME_HIP(hipStreamSynchronize(s));
      st->batches++;
    }
  }
  uint64_t* pin = ctx->pinned();
  SME_HIP(hipMemcpyAsync(pin, acc.get(), 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SME_HIP(hipStreamSynchronize(s));
  st->d2h_syncs++;
  (void)wedges;
  *count = pin[0] - 2 * pin[1];
}


// Diamond (Q2, LIGHT, ordering on) over all start vertices: every level
// output is a whole-graph count -- start vertices n, edges m, triangles
// through an edge summed over edges (3T), and the count D = sum_e
// t(e)(t(e)-1)/2 (two triangles sharing the chord, LIGHTDirectCount.cuh:
// 31-37) -- so it follows from the per-edge triangle counts of any
// isomorphic copy.
bool LocalDiamondApplies(const Context* ctx, const DevGraph* g, const EnginePlan* plan) {
  return !ctx->tuning.census_level_engine && plan->algo == kAlgoLight && plan->query_type == 2 &&
         plan->enable_ordering && !g->self_loops && g->max_up_degree <= kLocalMax &&
         CensusBytes(g) <= ctx->RunAvailable();
}

void LocalDiamond(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint64_t* count,
                  uint64_t* levels, RunStats* st) {
  hipStream_t s = ctx->stream();
  const uint64_t n = g->n, nnz = g->nnz;
  *st = RunStats{};
  st->batches = 1;
  uint64_t tri = 0, tri_levels[kMaxOps] = {0};
  Buf<uint32_t> t(ctx, nnz);
  Buf<unsigned long long> acc(ctx, 2);
  LaunchFill64(reinterpret_cast<uint64_t*>(acc.get()), 2, 0, s);
  if (nnz) {
    LaunchFill32(t.get(), nnz, 0, s);
    EnginePlan tp;
    if (!MakeLightEnginePlan(0, kClSpecial, true, &tp, nullptr)) Fail(kErrInvalid, "triangle plan");
    RunStats rs;
    LocalClique(ctx, g, &tp, 0, static_cast<uint32_t>(n), BucketRef{}, &tri, tri_levels, &rs, t.get());
    st->d2h_syncs += rs.d2h_syncs;
    HouseArgs h{};
    h.rp = g->rp;
    h.up = g->up;
    h.n = n;
    h.t = t.get();
    h.acc = acc.get();
    Timed tm(ctx, s, kSlotHouseMirror, 4.0 * static_cast<double>(nnz) / 2 + 16.0 * static_cast<double>(n), nnz / 2);
    LaunchEdgePairs(h, s);
  }
  uint64_t* pin = ctx->pinned();
  SME_HIP(hipMemcpyAsync(pin, acc.get(), 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SME_HIP(hipStreamSynchronize(s));
  st->d2h_syncs++;
  const uint64_t three_t = pin[0], d = pin[1];
  if (three_t != 3 * tri) Fail(kErrInvalid, "per-edge triangle counts disagree with the triangle count");
  *count = d;
  if (levels) {
    // tri_levels of the triangle plan (M0 C1 M1 CC2): n, m, m, T
    const uint64_t m = tri_levels[1];
    for (int L = 0; L < plan->nops; ++L) {
      const OpDesc& o = plan->ops[L];
      uint64_t v = d;  // FILTER_COMPUTE, COMPUTE_COUNT, COUNT: pairs of triangles
      if (o.kind == kMaterialize && o.u == 0)
        v = n;
      else if ((o.kind == kCompute || o.kind == kMaterialize) && o.u == 2)
        v = m;
      else if (o.kind == kCompute || o.kind == kComputePathCount || o.kind == kMaterialize)
        v = three_t;
      levels[L] = v;
    }
  }
}

}  // namespace

Status RunPlan(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t v_begin,
               uint32_t v_end, uint64_t* count, uint64_t* level_counts, RunStats* stats,
               std::string* err, bool ranked_ids) {
  if (!ctx || !g || !plan || !count) {
    *err = "null argument";
    return kErrInvalid;
  }
  if (ranked_ids && !g->ranked) {
    *err = "the graph has no degree-ranked copy (upload it with rank order on)";
    return kErrInvalid;
  }
  if (g->device != ctx->device()) {
    *err = "graph lives on another device";
    return kErrInvalid;
  }
  if (v_end > g->n) v_end = static_cast<uint32_t>(g->n);
  if (v_begin > v_end) v_begin = v_end;
  // A count over all start vertices is the same on any isomorphic copy of
  // the graph, and so are a clique plan's level outputs; the degree-ranked
  // copy walks far shorter lists (each instance is reached from its
  // lowest-degree vertex first).
  const bool use_ranked =
      ranked_ids || (g->ranked && v_begin == 0 && v_end == g->n && (plan->clique || !level_counts));
  const DevGraph* rg = use_ranked ? g->ranked : g;
  const auto t0 = std::chrono::steady_clock::now();
  try {
    SME_HIP(hipSetDevice(ctx->device()));
    ctx->held_peak = ctx->held;
    RunStats st;
    // the diamond's level outputs are whole-graph counts too: the ranked
    // copy serves a full-range run with them
    const DevGraph* dg = g->ranked ? g->ranked : g;
    const bool full = !ranked_ids && v_begin == 0 && v_end == g->n;
    bool on_ranked = use_ranked;
    // the edge census needs per-edge triangle counts and bit rows in memory;
    // when the budget cannot hold them the level engine runs instead
    auto census = [&](auto&& body) {
      try {
        body();
        return true;
      } catch (const Error& e) {
        if (e.status != kErrOom) throw;
        return false;
      }
    };
    if (full && !level_counts && LocalHouseApplies(ctx, rg, plan) &&
        census([&] { LocalHouse(ctx, rg, count, &st); })) {
    } else if (full && LocalDiamondApplies(ctx, dg, plan) &&
               census([&] { LocalDiamond(ctx, dg, plan, count, level_counts, &st); })) {
      on_ranked = dg != g;
    } else if (LocalCliqueApplies(ctx, plan)) {
      LocalClique(ctx, rg, plan, v_begin, v_end, BucketRef{}, count, level_counts, &st);
    } else {
      Runner r(ctx, rg, plan);
      r.Run(v_begin, v_end, count, level_counts, &st);
    }
    st.ranked = on_ranked ? 1 : 0;
    if (ctx->profiling) ctx->ResolveMarks();
    st.peak_pool_bytes = ctx->held_peak;
    st.elapsed_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
  } catch (const Error& e) {
    return Recover(ctx, e.status, e.msg, err);
  } catch (const std::bad_alloc&) {
    return Recover(ctx, kErrOom, "host memory exhausted", err);
  }
  return kOk;
}

Status ListMatches(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t v_begin,
                   uint32_t v_end, uint32_t* out, uint64_t max_rows, uint64_t* total, uint64_t* rows,
                   std::string* err) {
  if (!ctx || !g || !plan || !total || !rows || (max_rows && !out)) {
    *err = "null argument";
    return kErrInvalid;
  }
  if (plan->algo != kAlgoGpsm && plan->algo != kAlgoGpsmReuse) {
    *err = "match lists need a full-materialization plan (GPU_GPSM or GPU_GPSM_REUSE)";
    return kErrUnsupported;
  }
  if (g->device != ctx->device()) {
    *err = "graph lives on another device";
    return kErrInvalid;
  }
  if (v_end > g->n) v_end = static_cast<uint32_t>(g->n);
  try {
    SME_HIP(hipSetDevice(ctx->device()));
    ctx->held_peak = ctx->held;
    Runner r(ctx, g, plan);
    r.SetMatchOutput(out, max_rows);
    RunStats st;
    r.Run(v_begin, v_end, total, nullptr, &st);
    if (ctx->profiling) ctx->ResolveMarks();
    *rows = r.ListedRows();
  } catch (const Error& e) {
    return Recover(ctx, e.status, e.msg, err);
  } catch (const std::bad_alloc&) {
    return Recover(ctx, kErrOom, "host memory exhausted", err);
  }
  return kOk;
}

Status RunPlanBuckets(Context* ctx, const DevGraph* g, const EnginePlan* plan, const uint32_t* bounds,
                      uint64_t nbuckets, uint64_t* counts, uint64_t* count, uint64_t* level_counts,
                      RunStats* stats, std::string* err) {
  if (!ctx || !g || !plan || !count || !bounds || !counts || nbuckets == 0) {
    *err = "null argument";
    return kErrInvalid;
  }
  if (g->device != ctx->device()) {
    *err = "graph lives on another device";
    return kErrInvalid;
  }
  for (uint64_t k = 0; k < nbuckets; ++k) {
    if (bounds[k] > bounds[k + 1]) {
      *err = "bucket bounds must be non-decreasing";
      return kErrInvalid;
    }
  }
  if (bounds[nbuckets] > g->n) {
    *err = "bucket bounds exceed the vertex count";
    return kErrInvalid;
  }
  const auto t0 = std::chrono::steady_clock::now();
  try {
    SME_HIP(hipSetDevice(ctx->device()));
    ctx->held_peak = ctx->held;
    Buf<uint32_t> dbounds(ctx, nbuckets + 1);
    Buf<unsigned long long> dcounts(ctx, nbuckets);
    SME_HIP(hipMemcpyAsync(dbounds.get(), bounds, (nbuckets + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                           ctx->stream()));
    LaunchFill64(reinterpret_cast<uint64_t*>(dcounts.get()), nbuckets, 0, ctx->stream());
    // vertex -> bucket table over the covered range (4 bytes per vertex),
    // when it takes at most a quarter of what the run may use
    const uint64_t range = bounds[nbuckets] - bounds[0];
    const bool use_table = nbuckets > 1 && 4 * range <= ctx->RunAvailable() / 4;
    Buf<uint32_t> table(ctx, use_table ? range : 0);
    if (use_table) LaunchBucketTable(dbounds.get(), nbuckets, range, table.get(), ctx->stream());
    RunStats st;
    if (LocalCliqueApplies(ctx, plan)) {
      BucketRef bk{};
      bk.bounds = dbounds.get();
      bk.nbuckets = nbuckets;
      bk.counts = dcounts.get();
      bk.table = table.get();
      bk.base = bounds[0];
      LocalClique(ctx, g, plan, bounds[0], bounds[nbuckets], bk, count, level_counts, &st);
    } else {
      Runner r(ctx, g, plan);
      r.SetBuckets(dbounds.get(), nbuckets, dcounts.get(), table.get(), bounds[0]);
      r.Run(bounds[0], bounds[nbuckets], count, level_counts, &st);
    }
    SME_HIP(hipMemcpyAsync(counts, dcounts.get(), nbuckets * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           ctx->stream()));
    SME_HIP(hipStreamSynchronize(ctx->stream()));
    if (ctx->profiling) ctx->ResolveMarks();
    st.peak_pool_bytes = ctx->held_peak;
    st.elapsed_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
  } catch (const Error& e) {
    return Recover(ctx, e.status, e.msg, err);
  } catch (const std::bad_alloc&) {
    return Recover(ctx, kErrOom, "host memory exhausted", err);
  }
  return kOk;
}

}  // namespace sme
