// kernels.hip -- gfx950 kernels of the enumeration engine (see kernels.h).
//
// All kernels are integer-only.  Item kernels are wave64-centric: one wave
// owns a chunk of 64 consecutive items, so the validity of a chunk is one
// __ballot word and stable compaction needs only a popcount prefix.
#include "kernels.h"

namespace sme {

namespace {

constexpr int kThreads = 256;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr uint64_t kMaxGrid = 16384;
constexpr uint32_t kStageCap = kStageEntries;

using u64 = unsigned long long;

__device__ __forceinline__ uint32_t Gather(const ColSet& cs, int w, uint64_t p) {
  const uint32_t* col = cs.inst[w];
  return col ? col[p] : cs.iota[w] + static_cast<uint32_t>(p);
}

// First index in [lo, hi) with a[i] >= key (GPUBinSearch.cuh:18-31 semantics).
__device__ __forceinline__ uint64_t LowerBound(const uint32_t* a, uint64_t lo, uint64_t hi,
                                               uint32_t key) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool Contains(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t key) {
  const uint64_t pos = LowerBound(a, lo, hi, key);
  return pos < hi && a[pos] == key;
}

__device__ __forceinline__ uint64_t SegmentOf(const ItemArgs& a, uint64_t c, uint64_t t) {
  // largest s in [chunk_seg[c], chunk_seg[c+1]] with offs[s] - offs[0] <= t
  uint64_t lo = a.chunk_seg[c];
  uint64_t hi = (c + 1 < a.nchunks) ? a.chunk_seg[c + 1] : a.nseg - 1;
  const uint64_t key = a.offs[0] + t;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (a.offs[mid] <= key)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ u64 WaveSum(u64 v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ uint64_t SegIndex(const CandRef& c, uint64_t p) {
  return c.idx ? c.idx[p] : p;
}

constexpr u64 kNoBucket = ~0ull;

// Bucket of start vertex v: largest k with bounds[k] <= v.
__device__ __forceinline__ u64 BucketOf(const BucketRef& b, uint32_t v) {
  if (b.table) return b.table[v - b.base];
  uint64_t lo = 0, hi = b.nbuckets - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (b.bounds[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Per-wave running bucket sum (wave-uniform).  Bucket counts are added
// into it while the wave's keys stay on one bucket and written with one
// atomic when the bucket changes, so a wave issues about one atomic per
// bucket it touches instead of one per chunk (a few buckets would otherwise
// serialize every wave on the same counters).
struct BucketAcc {
  u64 key;
  u64 sum;
};

__device__ __forceinline__ void BucketFlush(const BucketRef& b, BucketAcc* acc) {
  if ((threadIdx.x & (kWave - 1)) == 0 && acc->sum && acc->key != kNoBucket)
    atomicAdd(b.counts + acc->key, acc->sum);
  acc->key = kNoBucket;
  acc->sum = 0;
}

// Adds each lane's value under its bucket key (keys non-decreasing across
// the wave; kNoBucket for lanes with nothing).  Must be reached by all 64
// lanes.  Runs before the wave's last bucket are written at once; the last
// one stays in the accumulator.
__device__ __forceinline__ void BucketAccAdd(const BucketRef& b, BucketAcc* acc, u64 key, u64 val) {
  const int lane = threadIdx.x & (kWave - 1);
  if (__ballot(key != kNoBucket && key != acc->key) == 0) {
    acc->sum += WaveSum(val);
    return;
  }
  BucketFlush(b, acc);
  u64 v = val;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const u64 y = __shfl_up(v, o);
    const u64 k2 = __shfl_up(key, o);
    if (lane >= o && k2 == key) v += y;
  }
  const u64 next = __shfl_down(key, 1);
  const bool run_end = lane == kWave - 1 || next != key;
  const u64 live = __ballot(key != kNoBucket);  // not 0: some key differed from acc
  const int last = kWave - 1 - __clzll(static_cast<long long>(live));
  const u64 last_key = __shfl(key, last);
  const u64 last_sum = __shfl(v, last);
  if (run_end && key != kNoBucket && key != last_key && v) atomicAdd(b.counts + key, v);
  acc->key = last_key;
  acc->sum = last_sum;
}

// ThreadCountWithComputedVertex (CheckConstraintsCommon.cuh:131-155): how
// many y in arr satisfy  elem (op) y.
__device__ __forceinline__ uint64_t CountWithComputed(const uint32_t* arr, uint64_t n, uint32_t elem,
                                                      uint8_t op) {
  if (n == 0) return 0;
  const uint64_t pos = LowerBound(arr, 0, n, elem);
  const bool eq = pos < n && arr[pos] == elem;
  if (op == 0) return n - pos - (eq ? 1 : 0);  // LESS_THAN
  if (op == 1) return pos;                     // LARGER_THAN
  return n - (eq ? 1 : 0);                     // NON_EQUAL
}

// CountWithComputed over the pair's set minus the excluded values present
// (bit k of `in` set: e_k lies in arr).  For !=: n' - [x in arr'];
// for < / >: the excluded values on the counted side drop out.
__device__ __forceinline__ uint64_t PairCount(const uint32_t* arr, uint64_t n, uint32_t x, uint8_t op,
                                              uint8_t nexcl, uint8_t in, uint32_t e0, uint32_t e1) {
  uint64_t v = CountWithComputed(arr, n, x, op);
  for (int k = 0; k < nexcl; ++k) {
    if (!((in >> k) & 1)) continue;
    const uint32_t e = k == 0 ? e0 : e1;
    if (op == 0)
      v -= e > x ? 1 : 0;
    else if (op == 1)
      v -= e < x ? 1 : 0;
    else
      v -= e == x ? 0 : 1;
  }
  return v;
}

__device__ __forceinline__ uint64_t GridStride() {
  return static_cast<uint64_t>(gridDim.x) * blockDim.x;
}

uint64_t GridFor(uint64_t n, uint64_t per_block, uint64_t cap = kMaxGrid) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (cap == 0 || cap > kMaxGrid) cap = kMaxGrid;
  return g > cap ? cap : g;
}

// --------------------------------------------------------------------------
// Set-up of one path of an intersection: the admissible window [lo, hi)
// from the < / > conditions, the pivot list and the pivot's window [*b, *e).
__device__ __forceinline__ void PrepPath(const PrepComputeArgs& a, uint64_t p, uint64_t* pb,
                                         uint64_t* pe, uint8_t* ppiv) {
  uint64_t lo = 0, hi = 0x100000000ull;
  for (int i = 0; i < a.ngt; ++i) {
    const uint64_t v = Gather(a.cs, a.gt_v[i], p) + 1ull;
    lo = v > lo ? v : lo;
  }
  for (int i = 0; i < a.nlt; ++i) {
    const uint64_t v = Gather(a.cs, a.lt_v[i], p);
    hi = v < hi ? v : hi;
  }
  // pivot: list of minimum length (ThreadChoosePivotLevel,
  // CheckConstraintsCommon.cuh:10-29; first minimum wins).  A reused
  // parent segment is short, so it is cut to the window before comparing.
  uint64_t best_b = 0, best_e = 0;
  uint8_t best = 0;
  bool have = false, windowed = false;
  if (a.has_reuse) {
    const uint64_t seg = SegIndex(a.reuse, p);
    uint64_t b = a.reuse.beg[seg], e = b + a.reuse.len[seg];
    if (lo >= hi) {
      e = b;
    } else {
      if (lo > 0) b = LowerBound(a.reuse.base, b, e, static_cast<uint32_t>(lo));
      if (hi <= 0xffffffffull) e = LowerBound(a.reuse.base, b, e, static_cast<uint32_t>(hi));
    }
    best = a.nb;
    best_b = b;
    best_e = e;
    have = windowed = true;
  }
  for (int j = 0; j < a.nb; ++j) {
    const uint32_t v = Gather(a.cs, a.bconn[j], p);
    const uint64_t b = a.rp[v], e = a.rp[v + 1];
    if (!have || e - b < best_e - best_b) {
      best = static_cast<uint8_t>(j);
      best_b = b;
      best_e = e;
      have = true;
      windowed = false;
    }
  }
  if (a.stream_ratio && a.has_reuse && a.nb == 1 && best == a.nb && best_e > best_b &&
      best_e - best_b <= kStageEntries) {
    // stream the adjacency list instead when it is at most stream_ratio
    // times the (windowed) reused segment
    const uint32_t v = Gather(a.cs, a.bconn[0], p);
    const uint64_t b = a.rp[v], e = a.rp[v + 1];
    if (e - b <= static_cast<uint64_t>(a.stream_ratio) * (best_e - best_b)) {
      best = 0;
      best_b = b;
      best_e = e;
      windowed = false;
    }
  }
  uint64_t b = best_b, e = best_e;
  if (!windowed) {
    if (lo >= hi) {
      e = b;
    } else {
      if (lo > 0) b = LowerBound(a.cols, b, e, static_cast<uint32_t>(lo));
      if (hi <= 0xffffffffull) e = LowerBound(a.cols, b, e, static_cast<uint32_t>(hi));
    }
  }
  *pb = b;
  *pe = e;
  *ppiv = best;
}

__global__ __launch_bounds__(kThreads) void PrepComputeKernel(PrepComputeArgs a) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < a.npaths;
       p += GridStride()) {
    if (a.group_cap && a.has_reuse && a.reuse.len[SegIndex(a.reuse, p)] <= a.group_cap) {
      a.size[p] = 0;  // counted by GroupCountKernel
      a.pstart[p] = 0;
      if (a.piv) a.piv[p] = 0;
      continue;
    }
    uint64_t b, e;
    uint8_t pv;
    PrepPath(a, p, &b, &e, &pv);
    uint64_t n = e - b;
    if (a.wg_min && n >= a.wg_min) {  // counted by a workgroup of its own
      const uint64_t k = atomicAdd(a.wg_count, 1ull);
      a.wg_list[2 * k] = p;
      a.wg_list[2 * k + 1] = n;
      n = 0;
    }
    a.size[p] = n;
    a.pstart[p] = b;
    if (a.piv) a.piv[p] = pv;
  }
}

__global__ __launch_bounds__(kThreads) void PrepListKernel(PrepListArgs a) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < a.npaths;
       p += GridStride()) {
    uint64_t lo = 0, hi = 0x100000000ull;
    for (int i = 0; i < a.ngt; ++i) {
      const uint64_t v = Gather(a.cs, a.gt_v[i], p) + 1ull;
      lo = v > lo ? v : lo;
    }
    for (int i = 0; i < a.nlt; ++i) {
      const uint64_t v = Gather(a.cs, a.lt_v[i], p);
      hi = v < hi ? v : hi;
    }
    const uint64_t seg = SegIndex(a.cand, p);
    uint64_t b = a.cand.beg[seg], e = b + a.cand.len[seg];
    if (lo >= hi) {
      e = b;
    } else {
      if (lo > 0) b = LowerBound(a.cand.base, b, e, static_cast<uint32_t>(lo));
      if (hi <= 0xffffffffull) e = LowerBound(a.cand.base, b, e, static_cast<uint32_t>(hi));
    }
    a.size[p] = e - b;
    a.pstart[p] = b;
  }
}

__global__ __launch_bounds__(kThreads) void PrepTwoKernel(PrepTwoArgs a) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < a.npaths;
       p += GridStride()) {
    const uint64_t s0 = SegIndex(a.a, p), s1 = SegIndex(a.b, p);
    const uint64_t n0 = a.a.len[s0], n1 = a.b.len[s1];
    const bool sw = !(n0 <= n1);  // GeneralCountTwoComputedVertices: iterate the smaller
    a.size[p] = sw ? n1 : n0;
    a.pstart[p] = sw ? a.b.beg[s1] : a.a.beg[s0];
    a.swapped[p] = sw ? 1 : 0;
  }
}

// ThreadDirectCount, LIGHTDirectCount.cuh:21-78.
__device__ uint64_t DirectCount(const uint64_t* c, int q) {
  uint64_t r = 0;
  switch (q) {
    case 0: r = c[2]; break;
    case 1: r = c[3]; break;
    case 2: if (c[1] > 0) r = c[1] * (c[1] - 1) / 2; break;
    case 3: r = c[3]; break;
    case 5:
      r = c[1] * c[3] * c[5];
      if (r > 0) r = r - c[3] * (c[1] - 1 + c[5] - 1) - c[3] * c[3];
      break;
    case 6: if (c[4] > 0 && c[2] > 0) r = c[4] * c[2] - c[2]; break;
    case 7: r = c[4]; break;
    case 8: r = c[4]; break;
    case 10: if (c[2] > 0) r = c[2] * (c[2] - 1) / 2; break;
    case 11: r = c[5]; break;
    default: r = 0; break;
  }
  return r;
}

__global__ __launch_bounds__(kThreads) void CountPathsKernel(CountPathArgs a) {
  u64 local = 0;
  const int lane = threadIdx.x & (kWave - 1);
  BucketAcc bacc{kNoBucket, 0};
  // wave-uniform trip count, so per-bucket sums can use the whole wave
  for (uint64_t base = blockIdx.x * static_cast<uint64_t>(blockDim.x) + (threadIdx.x - lane);
       base < a.npaths; base += GridStride()) {
    const uint64_t p = base + lane;
    u64 val = 0;
    if (p < a.npaths) {
      if (a.mode == 2) {
        val = 1;
      } else {
        uint64_t c[kQ] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < a.ncun; ++i) {
          const int u = a.cun[i];
          c[u] = a.cand[u].len[SegIndex(a.cand[u], p)];
        }
        val = a.mode == 0 ? DirectCount(c, a.query_type) : c[a.cun[0]];
      }
    }
    local += val;
    if (a.bk.bounds) {
      const u64 key = p < a.npaths ? BucketOf(a.bk, Gather(a.cs, a.bk.start_v, p)) : kNoBucket;
      BucketAccAdd(a.bk, &bacc, key, val);
    }
  }
  if (a.bk.bounds) BucketFlush(a.bk, &bacc);
  local = WaveSum(local);
  if ((threadIdx.x & (kWave - 1)) == 0 && local) {
    if (a.acc) atomicAdd(a.acc, local);
    if (a.acc_level) atomicAdd(a.acc_level, local);
  }
}

// chunk_seg[c] = segment of item 64c, written by the owning segment.
__global__ __launch_bounds__(kThreads) void ChunkPartitionKernel(const uint64_t* offs, uint64_t nseg,
                                                                 uint64_t nchunks, uint32_t* chunk_seg) {
  const uint64_t base = offs[0];
  for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < nseg;
       s += GridStride()) {
    const uint64_t b = offs[s] - base, e = offs[s + 1] - base;
    if (b == e) continue;
    uint64_t c = (b + kWave - 1) / kWave;
    const uint64_t last = (e - 1) / kWave;
    for (; c <= last && c < nchunks; ++c) chunk_seg[c] = static_cast<uint32_t>(s);
  }
}

// Membership + non-equality test of one item (the verify functor of
// GpsmBinSearch, GPSMBinSearch.cuh:146-170, with the < / > part already
// applied through the window).
__device__ __forceinline__ bool ItemValidPv(const ItemArgs& a, uint64_t s, uint8_t pv, uint32_t x) {
  if (a.nb + a.has_reuse > 1) {
    for (int j = 0; j < a.nb; ++j) {
      if (j == pv) continue;
      const uint32_t v = Gather(a.cs, a.bconn[j], s);
      const uint32_t h = a.hub_id ? a.hub_id[v] : kNoHub;
      if (h != kNoHub) {
        if (!((a.hub_bits[h * a.hub_words + (x >> 5)] >> (x & 31)) & 1u)) return false;
      } else if (!Contains(a.cols, a.rp[v], a.rp[v + 1], x)) {
        return false;
      }
    }
    if (a.has_reuse && pv != a.nb) {
      const uint64_t seg = SegIndex(a.reuse, s);
      const uint64_t b = a.reuse.beg[seg];
      if (!Contains(a.reuse.base, b, b + a.reuse.len[seg], x)) return false;
    }
  }
  for (int i = 0; i < a.nne; ++i)
    if (x == Gather(a.cs, a.ne_v[i], s)) return false;
  return true;
}

__device__ __forceinline__ bool ItemValid(const ItemArgs& a, uint64_t s, uint32_t x) {
  return ItemValidPv(a, s, a.nb + a.has_reuse > 1 ? a.piv[s] : 0, x);
}

// Values of segment s come from the graph's lists or, when the pivot is the
// reused parent segment, from that candidate buffer.
__device__ __forceinline__ const uint32_t* ItemSrc(const ItemArgs& a, uint64_t s) {
  return (a.has_reuse && a.piv[s] == a.nb) ? a.reuse.base : a.src;
}

// One 64-item chunk through the general per-lane path (items of several
// segments).  Returns the lane's count contribution (sum modes) or writes
// the chunk's mask and popcount (mask mode).
template <int kMode>
__device__ __forceinline__ u64 GenericChunk(const ItemArgs& a, uint64_t c, int lane, BucketAcc* bacc) {
  const uint64_t t = c * kWave + lane;
  bool ok = t < a.nitems;
  uint64_t s = 0;
  u64 val = 0;
  if (ok) {
    s = SegmentOf(a, c, t);
    const uint64_t r = a.offs[0] + t - a.offs[s];
    if (kMode == kModeSumTwo) {
      // GeneralCountTwoComputedVertices, LIGHTCount.cuh:22-116: walk the
      // smaller set, binary-search the larger one.
      const bool sw = a.swapped[s] != 0;
      const CandRef& sm = sw ? a.second : a.pair;
      const CandRef& bg = sw ? a.pair : a.second;
      const uint32_t x = sm.base[a.pstart[s] + r];
      const uint64_t bs = SegIndex(bg, s);
      const uint32_t* arr = bg.base + bg.beg[bs];
      const uint64_t n = bg.len[bs];
      const uint64_t pos = LowerBound(arr, 0, n, x);
      const bool eq = pos < n && arr[pos] == x;
      if (a.pair_op == 0)
        val = sw ? pos : n - pos - (eq ? 1 : 0);
      else if (a.pair_op == 1)
        val = sw ? n - pos - (eq ? 1 : 0) : pos;
      else
        val = n - (eq ? 1 : 0);
    } else {
      const uint32_t x = ItemSrc(a, s)[a.pstart[s] + r];
      ok = ItemValid(a, s, x);
      if (kMode == kModeSum) {
        val = ok ? 1 : 0;
      } else if (kMode == kModeSumPair && ok) {
        const uint64_t ps = SegIndex(a.pair, s);
        if (a.pair_nexcl) {
          val = PairCount(a.pair.base + a.pair.beg[ps], a.pair.len[ps], x, a.pair_op, a.pair_nexcl,
                          a.pair_excl_in[s], Gather(a.cs, a.pair_excl_v[0], s),
                          a.pair_nexcl > 1 ? Gather(a.cs, a.pair_excl_v[1], s) : 0u);
        } else {
          val = CountWithComputed(a.pair.base + a.pair.beg[ps], a.pair.len[ps], x, a.pair_op);
        }
      }
    }
  }
  if (kMode == kModeMask) {
    const u64 m = __ballot(ok);
    if (lane == 0) {
      a.mask[c] = m;
      a.ccount[c] = static_cast<uint32_t>(__popcll(m));
    }
  } else if (a.bk.bounds) {
    const u64 key = t < a.nitems ? BucketOf(a.bk, Gather(a.cs, a.bk.start_v, s)) : kNoBucket;
    BucketAccAdd(a.bk, bacc, key, val);
  }
  return val;
}

// Wave-level ordering point for LDS staging: every lane's earlier LDS
// accesses complete before any lane's later ones.
__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t Uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ bool ContainsLds(const uint32_t* a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && a[lo] == key;
}

// A sorted list to test against: a hub bitmap row, or entries a[lo, hi).
struct ListRef {
  const uint32_t* a;
  const uint32_t* hrow;
  uint64_t lo, hi;
};

// Lists other than the pivot of segment s (at most two are kept; callers
// check OtherLists(a) <= 2 first).
__device__ __forceinline__ int OtherLists(const ItemArgs& a) {
  return a.nb + a.has_reuse > 1 ? a.nb + a.has_reuse - 1 : 0;
}

// The k-th list (k = 0, 1) other than the pivot pv of segment s.
__device__ __forceinline__ ListRef OtherList(const ItemArgs& a, uint64_t s, uint8_t pv, int k) {
  // adjacency lists in bconn order skipping the pivot, then the reused segment
  int j = k + (pv <= k ? 1 : 0);
  if (pv == a.nb) j = k;  // the pivot is the reused segment
  if (j < a.nb) {
    const uint32_t v = Gather(a.cs, a.bconn[j], s);
    const uint32_t h = a.hub_id ? a.hub_id[v] : kNoHub;
    return ListRef{a.cols, h != kNoHub ? a.hub_bits + static_cast<uint64_t>(h) * a.hub_words : nullptr,
                   a.rp[v], a.rp[v + 1]};
  }
  const uint64_t seg = SegIndex(a.reuse, s);
  const uint64_t b = a.reuse.beg[seg];
  return ListRef{a.reuse.base, nullptr, b, b + a.reuse.len[seg]};
}

__device__ __forceinline__ void CollectLists(const ItemArgs& a, uint64_t s, uint8_t pv, ListRef* l0,
                                             ListRef* l1) {
  const int n = OtherLists(a);
  if (n > 0) *l0 = OtherList(a, s, pv, 0);
  if (n > 1) *l1 = OtherList(a, s, pv, 1);
}

__device__ __forceinline__ void NarrowToWindow(ListRef* r, uint32_t x_first, uint32_t x_last) {
  const uint64_t lo = LowerBound(r->a, r->lo, r->hi, x_first);
  uint64_t hi = LowerBound(r->a, lo, r->hi, x_last);
  if (hi < r->hi && r->a[hi] == x_last) ++hi;
  r->lo = lo;
  r->hi = hi;
}

__device__ __forceinline__ bool Member(const ListRef& r, bool staged, const uint32_t* stage,
                                       uint32_t nstage, uint32_t x) {
  if (r.hrow) return (r.hrow[x >> 5] >> (x & 31)) & 1u;
  if (staged) return ContainsLds(stage, nstage, x);
  return Contains(r.a, r.lo, r.hi, x);
}

constexpr int kProbeChunks = 4;

// Membership of x[j] in the sorted a[0, n) for kProbeChunks keys at once:
// one bisection for all of them (its steps depend on n only), each step
// issuing the keys' loads together.  The invariant is that the key's lower
// bound lies in [b, b + len]; at the end it is b or b + 1.
__device__ __forceinline__ void SearchN(const uint32_t* a, uint32_t n, const uint32_t* x, bool* ok) {
  uint32_t b[kProbeChunks];
#pragma unroll
  for (int j = 0; j < kProbeChunks; ++j) b[j] = 0;
  uint32_t len = n;
  while (len > 1) {
    const uint32_t half = len >> 1;
#pragma unroll
    for (int j = 0; j < kProbeChunks; ++j) b[j] = a[b[j] + half - 1] < x[j] ? b[j] + half : b[j];
    len -= half;
  }
#pragma unroll
  for (int j = 0; j < kProbeChunks; ++j) {
    const bool at_b = b[j] < n && a[b[j]] == x[j];
    const bool at_b1 = b[j] + 1 < n && a[b[j] + 1] == x[j];
    ok[j] = ok[j] && (at_b || at_b1);
  }
}

// Member() for kProbeChunks keys; a key whose ok is false is left false.
__device__ __forceinline__ void MemberN(const ListRef& r, bool staged, const uint32_t* stage,
                                        uint32_t nstage, const uint32_t* x, bool* ok) {
  if (r.hrow) {
    uint32_t w[kProbeChunks];
#pragma unroll
    for (int j = 0; j < kProbeChunks; ++j) w[j] = ok[j] ? r.hrow[x[j] >> 5] : 0u;
#pragma unroll
    for (int j = 0; j < kProbeChunks; ++j) ok[j] = ok[j] && ((w[j] >> (x[j] & 31)) & 1u);
  } else if (staged) {
    SearchN(stage, nstage, x, ok);
  } else {
    SearchN(r.a + r.lo, static_cast<uint32_t>(r.hi - r.lo), x, ok);
  }
}

// Chunks [c0, c1) whose items all belong to segment s (at most two other
// lists and two non-equality conditions).  Everything per segment is loaded
// once.  Every list is cut to the value window [x_first, x_last] of the run
// (the items are one sorted run of the pivot); the first non-hub list whose
// window has at most kStageCap entries is staged in LDS (independent loads:
// one memory round trip), turning its membership tests into LDS binary
// searches.  A list that is short already is staged whole, without the two
// searches for its window.
template <int kMode>
__device__ u64 FastSegment(const ItemArgs& a, uint64_t s, uint64_t c0, uint64_t c1,
                           uint32_t* stage, int lane) {
  s = Uniform64(s);
  const uint8_t pv = (a.nb + a.has_reuse > 1) ? a.piv[s] : 0;
  const uint32_t* src = (a.has_reuse && pv == a.nb) ? a.reuse.base : a.src;
  const uint64_t t0 = c0 * kWave;
  const uint64_t t1 = (c1 * kWave < a.nitems) ? c1 * kWave : a.nitems;
  const uint64_t base = a.pstart[s] + (a.offs[0] + t0 - a.offs[s]);
  const int nl = OtherLists(a);
  ListRef l0{nullptr, nullptr, 0, 0}, l1{nullptr, nullptr, 0, 0};
  CollectLists(a, s, pv, &l0, &l1);
  // staged list: 0 = l0, 1 = l1, -1 none
  int staged = -1;
  const uint32_t x_first = src[base], x_last = src[base + (t1 - t0) - 1];
  bool n0 = false, n1 = false;
  if (nl > 0 && !l0.hrow) {
    if (l0.hi - l0.lo > kStageCap) {
      NarrowToWindow(&l0, x_first, x_last);
      n0 = true;
    }
    if (l0.hi - l0.lo <= kStageCap) staged = 0;
  }
  if (staged < 0 && nl > 1 && !l1.hrow) {
    if (l1.hi - l1.lo > kStageCap) {
      NarrowToWindow(&l1, x_first, x_last);
      n1 = true;
    }
    if (l1.hi - l1.lo <= kStageCap) staged = 1;
  }
  uint32_t nstage = 0;
  if (staged >= 0) {
    const uint32_t* sa = staged == 0 ? l0.a : l1.a;
    const uint64_t slo = staged == 0 ? l0.lo : l1.lo;
    nstage = static_cast<uint32_t>((staged == 0 ? l0.hi : l1.hi) - slo);
    WaveSync();  // earlier readers of the stage are done
    for (uint32_t i = lane; i < nstage; i += kWave) stage[i] = sa[slo + i];
    WaveSync();
  }
  if (nl > 0 && !l0.hrow && staged != 0 && !n0) NarrowToWindow(&l0, x_first, x_last);
  if (nl > 1 && !l1.hrow && staged != 1 && !n1) NarrowToWindow(&l1, x_first, x_last);
  const uint32_t ne0 = a.nne > 0 ? Gather(a.cs, a.ne_v[0], s) : 0u;
  const uint32_t ne1 = a.nne > 1 ? Gather(a.cs, a.ne_v[1], s) : 0u;
  const uint32_t* parr = nullptr;
  uint64_t plen = 0;
  uint8_t pin = 0;
  uint32_t pe0 = 0, pe1 = 0;
  if (kMode == kModeSumPair) {
    const uint64_t ps = SegIndex(a.pair, s);
    parr = a.pair.base + a.pair.beg[ps];
    plen = a.pair.len[ps];
    if (a.pair_nexcl) {
      pin = a.pair_excl_in[s];
      pe0 = Gather(a.cs, a.pair_excl_v[0], s);
      if (a.pair_nexcl > 1) pe1 = Gather(a.cs, a.pair_excl_v[1], s);
    }
  }
  // each lane takes one item of each of kProbeChunks chunks at a time, so
  // their membership probes are independent loads in flight together
  u64 local = 0;
  for (uint64_t cb = c0; cb < c1; cb += kProbeChunks) {
    uint32_t x[kProbeChunks];
    bool ok[kProbeChunks];
#pragma unroll
    for (int j = 0; j < kProbeChunks; ++j) {
      const uint64_t t = (cb + j) * kWave + lane;
      ok[j] = cb + j < c1 && t < t1;
      x[j] = ok[j] ? src[base + (t - t0)] : 0u;
    }
    if (nl > 0) MemberN(l0, staged == 0, stage, nstage, x, ok);
    if (nl > 1) MemberN(l1, staged == 1, stage, nstage, x, ok);
#pragma unroll
    for (int j = 0; j < kProbeChunks; ++j) {
      if (a.nne > 0) ok[j] = ok[j] && x[j] != ne0;
      if (a.nne > 1) ok[j] = ok[j] && x[j] != ne1;
      if (cb + j >= c1) continue;  // wave-uniform
      if (kMode == kModeMask) {
        const u64 m = __ballot(ok[j]);
        if (lane == 0) {
          a.mask[cb + j] = m;
          a.ccount[cb + j] = static_cast<uint32_t>(__popcll(m));
        }
      } else if (kMode == kModeSum) {
        local += ok[j] ? 1 : 0;
      } else if (ok[j]) {
        local += PairCount(parr, plen, x[j], a.pair_op, a.pair_nexcl, pin, pe0, pe1);
      }
    }
  }
  return local;
}

template <int kMode>
__global__ __launch_bounds__(kThreads) void VerifyKernel(ItemArgs a) {
  __shared__ uint32_t stage_all[kWavesPerBlock][kStageCap];
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t* stage = stage_all[threadIdx.x / kWave];
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = GridStride() / kWave;
  const uint64_t M = a.super_chunks > 1 ? a.super_chunks : 1;
  const uint64_t nsuper = (a.nchunks + M - 1) / M;
  u64 local = 0;
  BucketAcc bacc{kNoBucket, 0};
  for (uint64_t sc = wave; sc < nsuper; sc += nwaves) {
    const uint64_t c0 = sc * M;
    const uint64_t c1 = c0 + M < a.nchunks ? c0 + M : a.nchunks;
    bool fast = false;
    uint64_t s = 0;
    if (kMode != kModeSumTwo && M > 1 && OtherLists(a) <= 2 && a.nne <= 2) {
      const uint64_t t_last = (c1 * kWave < a.nitems ? c1 * kWave : a.nitems) - 1;
      s = a.chunk_seg[c0];
      fast = SegmentOf(a, c1 - 1, t_last) == s;
    }
    if (fast) {
      const u64 seg = FastSegment<kMode>(a, s, c0, c1, stage, lane);
      local += seg;
      if (kMode != kModeMask && a.bk.bounds)
        BucketAccAdd(a.bk, &bacc, BucketOf(a.bk, Gather(a.cs, a.bk.start_v, s)), seg);
    } else {
      for (uint64_t c = c0; c < c1; ++c) local += GenericChunk<kMode>(a, c, lane, &bacc);
    }
  }
  if (kMode != kModeMask && a.bk.bounds) BucketFlush(a.bk, &bacc);
  if (kMode != kModeMask) {
    local = WaveSum(local);
    if (lane == 0 && local) {
      atomicAdd(a.acc, local);
      atomicAdd(a.acc_level, local);
    }
  }
}

// Membership of x in the sorted list a[*pos, end) for increasing keys: the
// cursor only moves forward, and one round of 8 independent probes past it
// usually lands on the key's position; otherwise a binary search over the
// rest finds it.  Leaves *pos at the first entry >= x.
__device__ __forceinline__ bool Gallop(const uint32_t* a, uint64_t* pos, uint64_t end, uint32_t x) {
  const uint64_t p = *pos;
  uint32_t lt = 0;
  bool eq = false;
  const uint32_t* ap = a + p;
  const uint64_t room = end - p;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t y = static_cast<uint64_t>(k) < room ? ap[k] : 0xffffffffu;
    lt += y < x ? 1u : 0u;
    eq = eq || (y == x && static_cast<uint64_t>(k) < room);
  }
  if (lt < 8) {
    *pos = p + lt;
    return eq;
  }
  const uint64_t q = LowerBound(a, p + 8, end, x);
  *pos = q;
  return q < end && a[q] == x;
}

// Count of path p's pivot items [b, e) that pass every other list and the
// non-equality conditions (with the pair arithmetic for COMPUTE_COUNT with
// one other unmaterialized vertex).  One lane walks the path's items in
// order, galloping through each other list (at most two; with more, each
// item is binary-searched).
__device__ u64 LanePathCount(const PrepCountArgs& a, uint64_t p, uint8_t pv, uint64_t b,
                             uint64_t e) {
  const ItemArgs& it = a.it;
  const uint32_t* src = (it.has_reuse && pv == it.nb) ? it.reuse.base : it.cols;
  const uint32_t* parr = nullptr;
  uint64_t plen = 0;
  uint8_t pin = 0;
  uint32_t pe0 = 0, pe1 = 0;
  if (a.pair) {
    const uint64_t ps = SegIndex(it.pair, p);
    parr = it.pair.base + it.pair.beg[ps];
    plen = it.pair.len[ps];
    if (it.pair_nexcl) {
      pin = it.pair_excl_in[p];
      pe0 = Gather(it.cs, it.pair_excl_v[0], p);
      if (it.pair_nexcl > 1) pe1 = Gather(it.cs, it.pair_excl_v[1], p);
    }
  }
  u64 val = 0;
  const int nl = OtherLists(it);
  if (nl > 2 || it.nne > 2) {
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t x = src[i];
      if (ItemValidPv(it, p, pv, x))
        val += a.pair ? PairCount(parr, plen, x, it.pair_op, it.pair_nexcl, pin, pe0, pe1) : 1ull;
    }
    return val;
  }
  ListRef l0{nullptr, nullptr, 0, 0}, l1{nullptr, nullptr, 0, 0};
  CollectLists(it, p, pv, &l0, &l1);
  const uint32_t ne0 = it.nne > 0 ? Gather(it.cs, it.ne_v[0], p) : 0u;
  const uint32_t ne1 = it.nne > 1 ? Gather(it.cs, it.ne_v[1], p) : 0u;
  for (uint64_t i = b; i < e; ++i) {
    const uint32_t x = src[i];
    bool ok = true;
    if (nl > 0)
      ok = l0.hrow ? ((l0.hrow[x >> 5] >> (x & 31)) & 1u) != 0 : Gallop(l0.a, &l0.lo, l0.hi, x);
    if (nl > 1 && ok)
      ok = l1.hrow ? ((l1.hrow[x >> 5] >> (x & 31)) & 1u) != 0 : Gallop(l1.a, &l1.lo, l1.hi, x);
    if (it.nne > 0) ok = ok && x != ne0;
    if (it.nne > 1) ok = ok && x != ne1;
    if (ok) val += a.pair ? PairCount(parr, plen, x, it.pair_op, it.pair_nexcl, pin, pe0, pe1) : 1ull;
  }
  return val;
}

// COMPUTE_COUNT set-up that also counts short paths: a path whose pivot
// window holds at most lane_max items is verified by its own lane here and
// leaves size 0, so only longer paths go through the item schedule.  The
// loop is wave-uniform so per-bucket sums can use the whole wave.
__global__ __launch_bounds__(kThreads) void PrepCountKernel(PrepCountArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t np = a.pa.npaths;
  u64 local = 0, nlane = 0;
  BucketAcc bacc{kNoBucket, 0};
  for (uint64_t base = blockIdx.x * static_cast<uint64_t>(blockDim.x) + (threadIdx.x - lane);
       base < np; base += GridStride()) {
    const uint64_t p = base + lane;
    u64 val = 0;
    if (p < np) {
      uint64_t b, e;
      uint8_t pv;
      PrepPath(a.pa, p, &b, &e, &pv);
      uint64_t n = e - b;
      if (n <= a.lane_max) {
        if (n) val = LanePathCount(a, p, pv, b, e);
        nlane += n;
        n = 0;
      } else if (a.pa.wg_min && n >= a.pa.wg_min) {  // counted by a workgroup of its own
        const uint64_t k = atomicAdd(a.pa.wg_count, 1ull);
        a.pa.wg_list[2 * k] = p;
        a.pa.wg_list[2 * k + 1] = n;
        n = 0;
      }
      a.pa.size[p] = n;
      a.pa.pstart[p] = b;
      if (a.pa.piv) a.pa.piv[p] = pv;
    }
    local += val;
    if (a.it.bk.bounds) {
      const u64 key = p < np ? BucketOf(a.it.bk, Gather(a.it.cs, a.it.bk.start_v, p)) : kNoBucket;
      BucketAccAdd(a.it.bk, &bacc, key, val);
    }
  }
  if (a.it.bk.bounds) BucketFlush(a.it.bk, &bacc);
  local = WaveSum(local);
  nlane = WaveSum(nlane);
  if (lane == 0) {
    if (local) {
      atomicAdd(a.it.acc, local);
      atomicAdd(a.it.acc_level, local);
    }
    if (nlane && a.lane_items) atomicAdd(a.lane_items, nlane);
  }
}

__device__ __forceinline__ uint64_t RankOf(const RankRef& r, uint64_t t) {
  if (!r.mask) return t;
  const uint64_t c = t / kWave;
  if (c >= r.nchunks) return r.cbase[r.nchunks];
  const u64 below = (t & (kWave - 1)) ? (~0ull >> (kWave - (t & (kWave - 1)))) : 0ull;
  return r.cbase[c] + __popcll(r.mask[c] & below);
}

__global__ __launch_bounds__(kThreads) void CompactValuesKernel(CompactArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = GridStride() / kWave;
  const ItemArgs& it = a.it;
  for (uint64_t c = wave; c < it.nchunks; c += nwaves) {
    const uint64_t t = c * kWave + lane;
    if (t >= it.nitems) continue;
    const u64 m = a.rk.mask ? a.rk.mask[c] : ~0ull;
    if (!((m >> lane) & 1ull)) continue;
    const uint64_t s = SegmentOf(it, c, t);
    const uint64_t r = it.offs[0] + t - it.offs[s];
    const uint64_t pos = a.rk.mask ? a.rk.cbase[c] + __popcll(m & __lanemask_lt()) : t;
    a.out_vals[pos] = ItemSrc(it, s)[it.pstart[s] + r];
  }
}

__global__ __launch_bounds__(kThreads) void SegmentRanksKernel(CompactArgs a) {
  const ItemArgs& it = a.it;
  const uint64_t base = it.offs[0];
  for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < it.nseg;
       s += GridStride()) {
    const uint64_t b = RankOf(a.rk, it.offs[s] - base);
    const uint64_t e = RankOf(a.rk, it.offs[s + 1] - base);
    if (a.out_beg) a.out_beg[s] = b;
    a.out_len[s] = e - b;
  }
}

__global__ __launch_bounds__(kThreads) void ExpandKernel(ExpandArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = GridStride() / kWave;
  const ItemArgs& it = a.it;
  for (uint64_t c = wave; c < it.nchunks; c += nwaves) {
    const uint64_t t = c * kWave + lane;
    if (t >= it.nitems) continue;
    const u64 m = a.rk.mask ? a.rk.mask[c] : ~0ull;
    if (!((m >> lane) & 1ull)) continue;
    const uint64_t s = SegmentOf(it, c, t);
    const uint64_t r = it.offs[0] + t - it.offs[s];
    const uint64_t pos = a.rk.mask ? a.rk.cbase[c] + __popcll(m & __lanemask_lt()) : t;
    a.out_new[pos] = it.src[it.pstart[s] + r];
    for (int i = 0; i < a.ncarry; ++i) a.carry_out[i][pos] = Gather(it.cs, a.carry_v[i], s);
    for (int i = 0; i < a.nidx; ++i)
      a.idx_out[i][pos] = static_cast<uint32_t>(SegIndex(a.idx_in[i], s));
  }
}

// One work-group per hub: OR each neighbour's bit into the hub's row.
// Match-list output: row p of `out` is (M[0], ..., M[nv-1]) of path p.
__global__ __launch_bounds__(kThreads) void GatherTuplesKernel(ColSet cs, int nv, uint64_t n,
                                                               uint32_t* out) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n * nv;
       i += GridStride()) {
    const uint64_t p = i / nv;
    out[i] = Gather(cs, static_cast<int>(i - p * nv), p);
  }
}

__global__ __launch_bounds__(kThreads) void BucketTableKernel(BucketRef b, uint64_t range,
                                                              uint32_t* table) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < range;
       i += GridStride()) {
    table[i] = static_cast<uint32_t>(BucketOf(b, b.bounds[0] + static_cast<uint32_t>(i)));
  }
}

__global__ __launch_bounds__(kThreads) void ExclusionKernel(ExclusionArgs a) {
  u64 local = 0;
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < a.npaths;
       p += GridStride()) {
    const uint64_t seg = SegIndex(a.cand, p);
    const uint64_t b = a.cand.beg[seg], n = a.cand.len[seg];
    uint8_t bits = 0;
    uint64_t drop = 0;
    for (int k = 0; k < a.nne; ++k) {
      if (Contains(a.cand.base, b, b + n, Gather(a.cs, a.ne_v[k], p))) {
        bits |= static_cast<uint8_t>(1u << k);
        ++drop;
      }
    }
    // two equal excluded values count once
    if (a.nne == 2 && (bits & 3) == 3 && Gather(a.cs, a.ne_v[0], p) == Gather(a.cs, a.ne_v[1], p)) {
      bits = 1;
      drop = 1;
    }
    a.excl_in[p] = bits;
    local += n - drop;
  }
  local = WaveSum(local);
  if ((threadIdx.x & (kWave - 1)) == 0 && local) atomicAdd(a.acc_level, local);
}

__global__ __launch_bounds__(kThreads) void BuildHubBitsKernel(const uint64_t* rp, const uint32_t* cols,
                                                               const uint32_t* hubs, uint64_t nhub,
                                                               uint64_t words, uint32_t* bits) {
  for (uint64_t h = blockIdx.x; h < nhub; h += gridDim.x) {
    const uint32_t v = hubs[h];
    uint32_t* row = bits + h * words;
    for (uint64_t j = rp[v] + threadIdx.x; j < rp[v + 1]; j += blockDim.x) {
      const uint32_t x = cols[j];
      atomicOr(row + (x >> 5), 1u << (x & 31));
    }
  }
}

__global__ __launch_bounds__(kThreads) void Fill32Kernel(uint32_t* out, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += GridStride())
    out[i] = v;
}

__global__ __launch_bounds__(kThreads) void Iota32Kernel(uint32_t* out, uint64_t n, uint32_t base) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += GridStride())
    out[i] = base + static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(kThreads) void Fill64Kernel(uint64_t* out, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += GridStride())
    out[i] = v;
}

__global__ void SplitBatchesKernel(const uint64_t* offs, uint64_t start, uint64_t npaths, uint64_t bi,
                                   uint64_t bp, uint64_t budget, uint64_t max_items, uint64_t max_paths,
                                   uint64_t* bounds, uint64_t max_out, uint64_t* counters) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t k = 0;
  while (start < npaths && k < max_out) {
    // largest end in (start, limit] whose batch fits the budget and caps
    uint64_t lo = start + 1, hi = npaths;
    if (hi - start > max_paths) hi = start + max_paths;
    uint64_t best = start + 1;
    while (lo <= hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      const uint64_t items = offs[mid] - offs[start];
      const uint64_t cost = items * bi + (mid - start) * bp;
      if (cost <= budget && items <= max_items) {
        best = mid;
        lo = mid + 1;
      } else {
        hi = mid - 1;
      }
    }
    bounds[2 * k] = best;
    bounds[2 * k + 1] = offs[best];
    ++k;
    start = best;
  }
  counters[0] = k;
}

// ---------------------------------------------------------------------------
// Exclusive scan in three passes over 2048-element tiles (256 threads x 8).
constexpr int kScanVT = 8;
constexpr uint64_t kScanTile = kThreads * kScanVT;

template <typename T>
__global__ __launch_bounds__(kThreads) void ScanTileSumsKernel(const T* in, uint64_t n, u64* sums) {
  __shared__ u64 wsum[kWavesPerBlock];
  const uint64_t base = blockIdx.x * kScanTile;
  u64 v = 0;
#pragma unroll
  for (int k = 0; k < kScanVT; ++k) {
    const uint64_t i = base + static_cast<uint64_t>(k) * kThreads + threadIdx.x;
    if (i < n) v += static_cast<u64>(in[i]);
  }
  v = WaveSum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += wsum[w];
    sums[blockIdx.x] = t;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void ScanTileApplyKernel(const T* in, uint64_t n,
                                                                const u64* tile_prefix, u64* out) {
  __shared__ u64 wsum[kWavesPerBlock];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const uint64_t base = blockIdx.x * kScanTile + static_cast<uint64_t>(threadIdx.x) * kScanVT;
  u64 vals[kScanVT];
  u64 tot = 0;
#pragma unroll
  for (int k = 0; k < kScanVT; ++k) {
    const uint64_t i = base + k;
    vals[k] = i < n ? static_cast<u64>(in[i]) : 0ull;
    tot += vals[k];
  }
  // inclusive wave scan of thread totals
  u64 inc = tot;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const u64 y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == kWave - 1) wsum[wid] = inc;
  __syncthreads();
  u64 wpre = 0, btot = 0;
  for (int w = 0; w < kWavesPerBlock; ++w) {
    if (w < wid) wpre += wsum[w];
    btot += wsum[w];
  }
  const u64 tp = tile_prefix ? tile_prefix[blockIdx.x] : 0ull;
  u64 run = tp + wpre + inc - tot;
#pragma unroll
  for (int k = 0; k < kScanVT; ++k) {
    const uint64_t i = base + k;
    if (i < n) out[i] = run;
    run += vals[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tp + btot;
}

template <typename T>
void ScanImpl(const T* in, uint64_t n, uint64_t* out, uint64_t* tmp, hipStream_t stream) {
  const uint64_t tiles = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
  u64* o = reinterpret_cast<u64*>(out);
  if (tiles == 1) {
    hipLaunchKernelGGL(ScanTileApplyKernel<T>, dim3(1), dim3(kThreads), 0, stream, in, n, nullptr, o);
    return;
  }
  u64* sums = reinterpret_cast<u64*>(tmp);
  hipLaunchKernelGGL(ScanTileSumsKernel<T>, dim3(tiles), dim3(kThreads), 0, stream, in, n, sums);
  ScanImpl<unsigned long long>(sums, tiles, reinterpret_cast<uint64_t*>(sums),
                               tmp + tiles + 1, stream);
  hipLaunchKernelGGL(ScanTileApplyKernel<T>, dim3(tiles), dim3(kThreads), 0, stream, in, n, sums, o);
}

// --------------------------------------------------------------------------
// PathWorkgroupKernel: one workgroup per long path of a counting level (the
// paths the set-up kernel listed).  Per path, every list is cut to the value
// window of the path's pivot items, the first non-hub window of at most
// kWgStage entries is staged in LDS by the whole workgroup, and the 256
// threads walk the pivot items in order (coalesced), testing each against
// the staged window (LDS bisection), hub rows (one bit) or the remaining
// window (global bisection).  At most two lists besides the pivot and two
// != conditions (the caller checks).
template <int kMode>
__global__ __launch_bounds__(kThreads) void PathWorkgroupKernel(ItemArgs a, const uint64_t* wg_list,
                                                                const unsigned long long* wg_count,
                                                                unsigned long long* items) {
  __shared__ uint32_t stage[kWgStage];
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t np = *wg_count;
  u64 local = 0, walked = 0;
  BucketAcc bacc{kNoBucket, 0};
  for (uint64_t i = blockIdx.x; i < np; i += gridDim.x) {
    const uint64_t p = wg_list[2 * i], n = wg_list[2 * i + 1];
    const uint8_t pv = (a.nb + a.has_reuse > 1) ? a.piv[p] : 0;
    const uint32_t* src = (a.has_reuse && pv == a.nb) ? a.reuse.base : a.src;
    const uint64_t b = a.pstart[p];
    const int nl = OtherLists(a);
    ListRef l0{nullptr, nullptr, 0, 0}, l1{nullptr, nullptr, 0, 0};
    CollectLists(a, p, pv, &l0, &l1);
    const uint32_t x_first = src[b], x_last = src[b + n - 1];
    if (nl > 0 && !l0.hrow) NarrowToWindow(&l0, x_first, x_last);
    if (nl > 1 && !l1.hrow) NarrowToWindow(&l1, x_first, x_last);
    int staged = -1;
    if (nl > 0 && !l0.hrow && l0.hi - l0.lo <= kWgStage) staged = 0;
    else if (nl > 1 && !l1.hrow && l1.hi - l1.lo <= kWgStage) staged = 1;
    uint32_t nstage = 0;
    __syncthreads();  // the previous path's readers of the stage are done
    if (staged >= 0) {
      const ListRef& r = staged == 0 ? l0 : l1;
      nstage = static_cast<uint32_t>(r.hi - r.lo);
      for (uint32_t k = threadIdx.x; k < nstage; k += kThreads) stage[k] = r.a[r.lo + k];
    }
    __syncthreads();
    const uint32_t ne0 = a.nne > 0 ? Gather(a.cs, a.ne_v[0], p) : 0u;
    const uint32_t ne1 = a.nne > 1 ? Gather(a.cs, a.ne_v[1], p) : 0u;
    const uint32_t* parr = nullptr;
    uint64_t plen = 0;
    uint8_t pin = 0;
    uint32_t pe0 = 0, pe1 = 0;
    if (kMode == kModeSumPair) {
      const uint64_t ps = SegIndex(a.pair, p);
      parr = a.pair.base + a.pair.beg[ps];
      plen = a.pair.len[ps];
      if (a.pair_nexcl) {
        pin = a.pair_excl_in[p];
        pe0 = Gather(a.cs, a.pair_excl_v[0], p);
        if (a.pair_nexcl > 1) pe1 = Gather(a.cs, a.pair_excl_v[1], p);
      }
    }
    u64 v = 0;
    for (uint64_t k = threadIdx.x; k < n; k += kThreads) {
      const uint32_t x = src[b + k];
      bool ok = true;
      if (nl > 0) ok = Member(l0, staged == 0, stage, nstage, x);
      if (nl > 1 && ok) ok = Member(l1, staged == 1, stage, nstage, x);
      if (a.nne > 0) ok = ok && x != ne0;
      if (a.nne > 1) ok = ok && x != ne1;
      if (ok) v += kMode == kModeSum ? 1ull : PairCount(parr, plen, x, a.pair_op, a.pair_nexcl, pin, pe0, pe1);
    }
    local += v;
    if (threadIdx.x == 0) walked += n;
    if (a.bk.bounds) BucketAccAdd(a.bk, &bacc, BucketOf(a.bk, Gather(a.cs, a.bk.start_v, p)), v);
  }
  if (a.bk.bounds) BucketFlush(a.bk, &bacc);
  local = WaveSum(local);
  if (lane == 0) {
    if (local) {
      atomicAdd(a.acc, local);
      atomicAdd(a.acc_level, local);
    }
    if (items && walked) atomicAdd(items, walked);
  }
}

// --------------------------------------------------------------------------
// GroupCountKernel (see GroupCountArgs).
//
// One wave per run of paths sharing a segment S.  S is staged in LDS as a
// sorted array (set[0, n)) and a hash table of 4-slot buckets at set +
// kGroupMax (at most one value per bucket on average).  A value goes to the
// first free slot of its bucket, or of the next bucket when all four are
// taken; a lookup reads a whole bucket with one 16-byte LDS load and moves
// on only when that bucket is full and lacks the value, so lanes rarely
// loop.  The children's list windows are cut into tiles of 64 consecutive
// entries whose owner is wave-uniform, so per tile the owner's bounds are
// scalar and the entries are one coalesced load.
constexpr uint32_t kEmptySlot = 0xffffffffu;   // never a vertex id (n < 2^32 - 1)
constexpr uint64_t kGroupUnit = 256;           // paths per work unit of a wave
constexpr int kGroupUnroll = 8;                // tiles of 64 list entries in flight per wave
constexpr uint32_t kGroupLdsWords = 5 * kGroupMax;  // sorted values + buckets (<= 1 value per bucket on average)

struct alignas(16) Bucket4 {
  uint32_t s[4];
};

// Multiplicative hash on a 24-bit product (v_mul_u32_u24 is full rate, a
// 32-bit multiply is not): the high byte of x is folded into the low bits,
// then the top bits of the low 32 bits of the product with an odd 24-bit
// constant select the bucket.  __umul24 returns int, so the product is
// shifted as unsigned.
__device__ __forceinline__ uint32_t HashSlot(uint32_t x, uint32_t shift) {
  return static_cast<uint32_t>(__umul24(x ^ (x >> 24), 0x9E3779u)) >> shift;
}

__device__ __forceinline__ void BucketInsert(uint32_t* tab, uint32_t shift, uint32_t mask, uint32_t x) {
  uint32_t b = HashSlot(x, shift);
  while (true) {
    uint32_t* slot = tab + 4 * b;
    for (int k = 0; k < 4; ++k)
      if (atomicCAS(slot + k, kEmptySlot, x) == kEmptySlot) return;
    b = (b + 1) & mask;
  }
}

// Number of values of the sorted a[0, n) below key; the steps depend on n only.
__device__ __forceinline__ uint32_t RankLds(const uint32_t* a, uint32_t n, uint64_t key) {
  if (n == 0) return 0;
  uint32_t b = 0, len = n;
  while (len > 1) {
    const uint32_t half = len >> 1;
    b = a[b + half - 1] < key ? b + half : b;
    len -= half;
  }
  return b + (a[b] < key ? 1u : 0u);
}

__device__ __forceinline__ uint32_t ReadLane32(uint32_t v, uint32_t lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
}

__device__ __forceinline__ uint64_t ReadLane64(uint64_t v, uint32_t lane) {
  return (static_cast<uint64_t>(ReadLane32(static_cast<uint32_t>(v >> 32), lane)) << 32) |
         ReadLane32(static_cast<uint32_t>(v), lane);
}

__device__ __forceinline__ uint32_t Scalar32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <bool kNe>
__global__ __launch_bounds__(kThreads) void GroupCountKernel(GroupCountArgs a) {
  __shared__ uint32_t lds_set[kWavesPerBlock][kGroupLdsWords] __attribute__((aligned(16)));
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t* set = lds_set[threadIdx.x / kWave];
  const Bucket4* tab = reinterpret_cast<const Bucket4*>(set + kGroupMax);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = GridStride() / kWave;
  const uint64_t nunits = (a.npaths + kGroupUnit - 1) / kGroupUnit;
  u64 local = 0, walked = 0, lower = 0;
  BucketAcc bacc{kNoBucket, 0};
  for (uint64_t unit = wave; unit < nunits; unit += nwaves) {
    const uint64_t pb = unit * kGroupUnit;
    const uint64_t pe = pb + kGroupUnit < a.npaths ? pb + kGroupUnit : a.npaths;
    uint64_t staged = ~0ull;
    uint32_t n = 0, shift = 0, mask = 0, smin = 0, smax = 0;  // wave-uniform
    for (uint64_t sub = pb; sub < pe; sub += kWave) {
      const uint64_t p = sub + lane;
      const bool valid = p < pe;
      const uint64_t seg = valid ? SegIndex(a.seg, p) : ~0ull;
      u64 pending = __ballot(valid);
      while (pending) {
        // the run of paths sharing the lowest pending lane's segment (paths
        // of one segment are consecutive)
        const int f = __ffsll(static_cast<long long>(pending)) - 1;
        const uint64_t s = ReadLane64(seg, static_cast<uint32_t>(f));
        const bool in_run = valid && seg == s;
        pending &= ~__ballot(in_run);
        const uint64_t slen = a.seg.len[s];
        if (slen == 0 || slen > a.cap) continue;
        if (s != staged) {
          const uint32_t* src = a.seg.base + a.seg.beg[s];
          n = static_cast<uint32_t>(slen);
          uint32_t bbits = 3;  // buckets: a power of two >= n
          while ((1u << bbits) < n) ++bbits;
          shift = 32 - bbits;
          mask = (1u << bbits) - 1;
          WaveSync();  // earlier readers of the stage are done
          for (uint32_t i = lane; i < 4 * (mask + 1); i += kWave) set[kGroupMax + i] = kEmptySlot;
          WaveSync();
          for (uint32_t i = lane; i < n; i += kWave) {
            const uint32_t x = src[i];
            set[i] = x;
            BucketInsert(set + kGroupMax, shift, mask, x);
          }
          WaveSync();
          smin = Scalar32(set[0]);
          smax = Scalar32(set[n - 1]);
          staged = s;
        }
        // each child path's window of its adjacency list
        uint64_t wb = 0;
        uint32_t clen = 0, ne0 = kEmptySlot, ne1 = kEmptySlot;
        if (in_run) {
          const uint32_t v = Gather(a.cs, a.adj_v, p);
          uint64_t lo = smin, hi = static_cast<uint64_t>(smax) + 1;
          for (int i = 0; i < a.ngt; ++i) {
            const uint64_t g = Gather(a.cs, a.gt_v[i], p) + 1ull;
            lo = g > lo ? g : lo;
          }
          for (int i = 0; i < a.nlt; ++i) {
            const uint64_t l = Gather(a.cs, a.lt_v[i], p);
            hi = l < hi ? l : hi;
          }
          if (lo < hi) {
            const uint64_t rb = a.rp[v], re = a.rp[v + 1];
            const uint64_t b =
                (a.up && lo == static_cast<uint64_t>(v) + 1) ? a.up[v] : LowerBound(a.cols, rb, re, static_cast<uint32_t>(lo));
            // without a < condition the list is streamed to its end: values
            // above the segment's maximum fail the membership test
            const uint64_t e = a.nlt ? LowerBound(a.cols, b, re, static_cast<uint32_t>(hi)) : re;
            wb = b;
            clen = static_cast<uint32_t>(e - b);
          }
          if (a.nne > 0) ne0 = Gather(a.cs, a.ne_v[0], p);
          if (a.nne > 1) ne1 = Gather(a.cs, a.ne_v[1], p);
          if (a.lower && lo < hi) {
            const uint32_t sw = RankLds(set, n, hi) - RankLds(set, n, lo);
            lower += sw < clen ? sw : clen;
          }
        }
        // tiles of 64 consecutive entries: the owner of tile t is the
        // number of children whose tiles all lie before t, found by one
        // ballot, so per tile the owner's bounds are scalar.  In the run's
        // flat tile numbering, child c's entries are flat offsets
        // [64 excl_c, lend_c) and entry f is cols[abase_c + f].
        const uint32_t own_tiles = (clen + kWave - 1) / kWave;
        uint32_t incl = own_tiles;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
          const uint32_t t = __shfl_up(incl, o);
          if (lane >= o) incl += t;
        }
        const uint32_t excl = incl - own_tiles;
        const uint64_t abase = wb - static_cast<uint64_t>(excl) * kWave;  // wraps; only sums are used
        const uint32_t lend = excl * kWave + clen;
        const uint32_t ntiles = ReadLane32(incl, kWave - 1);
        u64 hits = 0;  // wave-uniform
        for (uint32_t tb = 0; tb < ntiles; tb += kGroupUnroll) {
          uint32_t y[kGroupUnroll];
          u64 okm[kGroupUnroll];
#pragma unroll
          for (int u = 0; u < kGroupUnroll; ++u) {  // issue every load first
            // a tile past the last repeats the last one with no lane valid
            const uint32_t tt = tb + u < ntiles ? tb + u : ntiles - 1;
            const uint32_t c = static_cast<uint32_t>(__popcll(__ballot(incl <= tt)));
            const uint32_t end = ReadLane32(lend, c);
            const uint32_t off = tt * kWave + lane;
            // a lane past its owner's window reads the owner's last entry
            y[u] = a.cols[ReadLane64(abase, c) + (off < end ? off : end - 1)];
            u64 m = tb + u < ntiles ? __ballot(off < end) : 0ull;
            if (kNe) {
              // the values tested here already passed the < / > window;
              // entries equal to a != operand of their path drop out
              const uint32_t x0 = ReadLane32(ne0, c), x1 = ReadLane32(ne1, c);
              m &= __ballot(y[u] != x0 && y[u] != x1);
            }
            okm[u] = m;
          }
          // one bucket load per entry; a valid entry is a hit when its
          // bucket holds it, and follows the chain when that bucket is full
          // (per lane, one bit per tile)
          uint32_t more = 0;
#pragma unroll
          for (int u = 0; u < kGroupUnroll; ++u) {
            const Bucket4 v = tab[HashSlot(y[u], shift)];
            const uint32_t d = min(min(v.s[0] ^ y[u], v.s[1] ^ y[u]), min(v.s[2] ^ y[u], v.s[3] ^ y[u]));
            hits += __popcll(okm[u] & __ballot(d == 0));
            const u64 chain = okm[u] & __ballot(d != 0 && v.s[3] != kEmptySlot);
            more |= ((chain >> lane) & 1ull) ? (1u << u) : 0u;
          }
          if (__ballot(more != 0)) {  // rare per entry: walk only the pending (lane, tile) pairs
            uint32_t extra = 0;
            while (more) {
              const int u = __ffs(more) - 1;
              more &= more - 1;
              uint32_t yy = y[0];
#pragma unroll
              for (int k = 1; k < kGroupUnroll; ++k) yy = u == k ? y[k] : yy;
              uint32_t b = HashSlot(yy, shift);
              bool go = true;
              while (go) {
                b = (b + 1) & mask;
                const Bucket4 v = tab[b];
                const bool hit = (v.s[0] == yy) || (v.s[1] == yy) || (v.s[2] == yy) || (v.s[3] == yy);
                extra += hit ? 1u : 0u;
                go = !hit && v.s[3] != kEmptySlot;
              }
            }
            hits += WaveSum(extra);
          }
        }
        if (a.walked) {
          const u64 run_len = WaveSum(clen);  // every lane takes part
          walked += lane == 0 ? run_len : 0;
        }
        local += lane == 0 ? hits : 0;
        if (a.bk.bounds) {
          const uint64_t first = sub + f;
          BucketAccAdd(a.bk, &bacc, BucketOf(a.bk, Gather(a.cs, a.bk.start_v, first)), lane == 0 ? hits : 0);
        }
      }
    }
  }
  if (a.bk.bounds) BucketFlush(a.bk, &bacc);
  local = WaveSum(local);
  lower = WaveSum(lower);
  walked = WaveSum(walked);
  if (lane == 0) {
    if (local) {
      atomicAdd(a.acc, local);
      atomicAdd(a.acc_level, local);
    }
    if (a.walked && walked) atomicAdd(a.walked, walked);
    if (a.lower && lower) atomicAdd(a.lower, lower);
  }
}

__global__ __launch_bounds__(kThreads) void UpperStartKernel(const uint64_t* rp, const uint32_t* cols,
                                                             uint64_t n, uint64_t* up) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += GridStride())
    up[v] = LowerBound(cols, rp[v], rp[v + 1], static_cast<uint32_t>(v + 1));
}

}  // namespace

size_t ScanTempWords(uint64_t n) {
  size_t words = 0;
  uint64_t m = n;
  while (true) {
    const uint64_t tiles = m == 0 ? 1 : (m + kScanTile - 1) / kScanTile;
    if (tiles == 1) break;
    words += tiles + 1;
    m = tiles;
  }
  return words + 1;
}

void LaunchScanU64(const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* tmp, hipStream_t stream) {
  ScanImpl<unsigned long long>(reinterpret_cast<const u64*>(in), n, out, tmp, stream);
}

void LaunchScanU32(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* tmp, hipStream_t stream) {
  ScanImpl<uint32_t>(in, n, out, tmp, stream);
}

void LaunchPrepCompute(const PrepComputeArgs& a, hipStream_t stream) {
  if (a.npaths == 0) return;
  hipLaunchKernelGGL(PrepComputeKernel, dim3(GridFor(a.npaths, kThreads)), dim3(kThreads), 0, stream, a);
}

void LaunchPrepCount(const PrepCountArgs& a, hipStream_t stream) {
  if (a.pa.npaths == 0) return;
  hipLaunchKernelGGL(PrepCountKernel, dim3(GridFor(a.pa.npaths, kThreads)), dim3(kThreads), 0, stream, a);
}

void LaunchExclusion(const ExclusionArgs& a, hipStream_t stream) {
  if (a.npaths == 0) return;
  hipLaunchKernelGGL(ExclusionKernel, dim3(GridFor(a.npaths, kThreads)), dim3(kThreads), 0, stream, a);
}

void LaunchPrepList(const PrepListArgs& a, hipStream_t stream) {
  if (a.npaths == 0) return;
  hipLaunchKernelGGL(PrepListKernel, dim3(GridFor(a.npaths, kThreads)), dim3(kThreads), 0, stream, a);
}

void LaunchPrepTwo(const PrepTwoArgs& a, hipStream_t stream) {
  if (a.npaths == 0) return;
  hipLaunchKernelGGL(PrepTwoKernel, dim3(GridFor(a.npaths, kThreads)), dim3(kThreads), 0, stream, a);
}

void LaunchCountPaths(const CountPathArgs& a, hipStream_t stream) {
  if (a.npaths == 0) return;
  hipLaunchKernelGGL(CountPathsKernel, dim3(GridFor(a.npaths, kThreads)), dim3(kThreads), 0, stream, a);
}

void LaunchChunkPartition(const uint64_t* offs, uint64_t nseg, uint64_t nchunks, uint32_t* chunk_seg,
                          hipStream_t stream) {
  if (nseg == 0 || nchunks == 0) return;
  hipLaunchKernelGGL(ChunkPartitionKernel, dim3(GridFor(nseg, kThreads)), dim3(kThreads), 0, stream, offs, nseg, nchunks, chunk_seg);
}

void LaunchVerify(const ItemArgs& a, int mode, uint32_t max_grid, hipStream_t stream) {
  if (a.nchunks == 0) return;
  const uint64_t M = a.super_chunks > 1 ? a.super_chunks : 1;
  const uint64_t grid = GridFor((a.nchunks + M - 1) / M, kWavesPerBlock, max_grid);
  switch (mode) {
    case kModeMask: hipLaunchKernelGGL(VerifyKernel<kModeMask>, dim3(grid), dim3(kThreads), 0, stream, a); break;
    case kModeSum: hipLaunchKernelGGL(VerifyKernel<kModeSum>, dim3(grid), dim3(kThreads), 0, stream, a); break;
    case kModeSumPair: hipLaunchKernelGGL(VerifyKernel<kModeSumPair>, dim3(grid), dim3(kThreads), 0, stream, a); break;
    case kModeSumTwo: hipLaunchKernelGGL(VerifyKernel<kModeSumTwo>, dim3(grid), dim3(kThreads), 0, stream, a); break;
  }
}

void LaunchCompact(const CompactArgs& a, hipStream_t stream) {
  if (a.out_vals && a.it.nchunks)
    hipLaunchKernelGGL(CompactValuesKernel, dim3(GridFor(a.it.nchunks, kWavesPerBlock)), dim3(kThreads), 0, stream, a);
  if (a.it.nseg) hipLaunchKernelGGL(SegmentRanksKernel, dim3(GridFor(a.it.nseg, kThreads)), dim3(kThreads), 0, stream, a);
}

void LaunchGroupCount(const GroupCountArgs& a, uint32_t max_grid, hipStream_t stream) {
  if (a.npaths == 0) return;
  const uint64_t units = (a.npaths + kGroupUnit - 1) / kGroupUnit;
  const dim3 grid(GridFor(units, kWavesPerBlock, max_grid));
  if (a.nne)
    hipLaunchKernelGGL(GroupCountKernel<true>, grid, dim3(kThreads), 0, stream, a);
  else
    hipLaunchKernelGGL(GroupCountKernel<false>, grid, dim3(kThreads), 0, stream, a);
}

void LaunchPathWorkgroup(const ItemArgs& a, int mode, const uint64_t* wg_list,
                         const unsigned long long* wg_count, unsigned long long* items,
                         uint32_t max_grid, hipStream_t stream) {
  // one workgroup per listed path; blocks beyond the list exit at once
  const dim3 grid(GridFor(static_cast<uint64_t>(max_grid ? max_grid : 2048), 1, max_grid));
  if (mode == kModeSumPair)
    hipLaunchKernelGGL(PathWorkgroupKernel<kModeSumPair>, grid, dim3(kThreads), 0, stream, a, wg_list,
                       wg_count, items);
  else
    hipLaunchKernelGGL(PathWorkgroupKernel<kModeSum>, grid, dim3(kThreads), 0, stream, a, wg_list,
                       wg_count, items);
}

void LaunchUpperStart(const uint64_t* rp, const uint32_t* cols, uint64_t n, uint64_t* up,
                      hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(UpperStartKernel, dim3(GridFor(n, kThreads)), dim3(kThreads), 0, stream, rp, cols, n, up);
}

void LaunchExpand(const ExpandArgs& a, hipStream_t stream) {
  if (a.it.nchunks == 0) return;
  hipLaunchKernelGGL(ExpandKernel, dim3(GridFor(a.it.nchunks, kWavesPerBlock)), dim3(kThreads), 0, stream, a);
}

void LaunchBuildHubBits(const uint64_t* rp, const uint32_t* cols, const uint32_t* hubs, uint64_t nhub,
                        uint64_t words, uint32_t* bits, hipStream_t stream) {
  if (nhub == 0) return;
  hipLaunchKernelGGL(BuildHubBitsKernel, dim3(GridFor(nhub, 1)), dim3(kThreads), 0, stream, rp, cols, hubs,
                     nhub, words, bits);
}

void LaunchGatherTuples(const ColSet& cs, int nv, uint64_t n, uint32_t* out, hipStream_t stream) {
  if (n == 0 || nv <= 0) return;
  hipLaunchKernelGGL(GatherTuplesKernel, dim3(GridFor(n * nv, kThreads)), dim3(kThreads), 0, stream, cs, nv,
                     n, out);
}

void LaunchBucketTable(const uint32_t* bounds, uint64_t nbuckets, uint64_t range, uint32_t* table,
                       hipStream_t stream) {
  if (range == 0) return;
  BucketRef b{};
  b.bounds = bounds;
  b.nbuckets = nbuckets;
  hipLaunchKernelGGL(BucketTableKernel, dim3(GridFor(range, kThreads)), dim3(kThreads), 0, stream, b, range,
                     table);
}

void LaunchFill32(uint32_t* out, uint64_t n, uint32_t value, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(Fill32Kernel, dim3(GridFor(n, kThreads)), dim3(kThreads), 0, stream, out, n, value);
}

void LaunchIota32(uint32_t* out, uint64_t n, uint32_t base, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(Iota32Kernel, dim3(GridFor(n, kThreads)), dim3(kThreads), 0, stream, out, n, base);
}

void LaunchFill64(uint64_t* out, uint64_t n, uint64_t value, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(Fill64Kernel, dim3(GridFor(n, kThreads)), dim3(kThreads), 0, stream, out, n, value);
}

void LaunchSplitBatches(const uint64_t* offs, uint64_t start, uint64_t npaths, uint64_t bi, uint64_t bp,
                        uint64_t budget, uint64_t max_items, uint64_t max_paths, uint64_t* bounds,
                        uint64_t max_out, uint64_t* counters, hipStream_t stream) {
  hipLaunchKernelGGL(SplitBatchesKernel, dim3(1), dim3(64), 0, stream, offs, start, npaths, bi, bp, budget, max_items, max_paths,
                                           bounds, max_out, counters);
}

}  // namespace sme
