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

// Segmented wave sum over lanes with equal keys (keys are non-decreasing
// across the wave); the last lane of each run adds the run's total.  Must
// be reached by all 64 lanes.
__device__ __forceinline__ void BucketAdd(const BucketRef& b, u64 key, u64 v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const u64 y = __shfl_up(v, o);
    const u64 k2 = __shfl_up(key, o);
    if (lane >= o && k2 == key) v += y;
  }
  const u64 next = __shfl_down(key, 1);
  const bool run_end = lane == kWave - 1 || next != key;
  if (run_end && v && key != kNoBucket) atomicAdd(b.counts + key, v);
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
    uint64_t b, e;
    uint8_t pv;
    PrepPath(a, p, &b, &e, &pv);
    a.size[p] = e - b;
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
      BucketAdd(a.bk, key, val);
    }
  }
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

template <int kMode>
__global__ __launch_bounds__(kThreads) void VerifyKernel(ItemArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = GridStride() / kWave;
  u64 local = 0;
  for (uint64_t c = wave; c < a.nchunks; c += nwaves) {
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
          val = CountWithComputed(a.pair.base + a.pair.beg[ps], a.pair.len[ps], x, a.pair_op);
        }
      }
    }
    if (kMode == kModeMask) {
      const u64 m = __ballot(ok);
      if (lane == 0) {
        a.mask[c] = m;
        a.ccount[c] = static_cast<uint32_t>(__popcll(m));
      }
    } else {
      local += val;
      if (a.bk.bounds) {
        const u64 key = t < a.nitems ? BucketOf(a.bk, Gather(a.cs, a.bk.start_v, s)) : kNoBucket;
        BucketAdd(a.bk, key, val);
      }
    }
  }
  if (kMode != kModeMask) {
    local = WaveSum(local);
    if (lane == 0 && local) {
      atomicAdd(a.acc, local);
      atomicAdd(a.acc_level, local);
    }
  }
}

// COMPUTE_COUNT set-up that also counts short paths: a path whose pivot
// window holds at most lane_max items is verified by its own lane here and
// leaves size 0, so only longer paths go through the item schedule.  The
// loop is wave-uniform so per-bucket sums can use the whole wave.
__global__ __launch_bounds__(kThreads) void PrepCountKernel(PrepCountArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t np = a.pa.npaths;
  u64 local = 0, nlane = 0;
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
        const uint32_t* src = (a.it.has_reuse && pv == a.it.nb) ? a.it.reuse.base : a.it.cols;
        uint64_t ps = 0, plen = 0;
        const uint32_t* parr = nullptr;
        if (a.pair) {
          ps = SegIndex(a.it.pair, p);
          parr = a.it.pair.base + a.it.pair.beg[ps];
          plen = a.it.pair.len[ps];
        }
        for (uint64_t i = b; i < e; ++i) {
          const uint32_t x = src[i];
          if (!ItemValidPv(a.it, p, pv, x)) continue;
          val += a.pair ? CountWithComputed(parr, plen, x, a.it.pair_op) : 1ull;
        }
        nlane += n;
        n = 0;
      }
      a.pa.size[p] = n;
      a.pa.pstart[p] = b;
      if (a.pa.piv) a.pa.piv[p] = pv;
    }
    local += val;
    if (a.it.bk.bounds) {
      const u64 key = p < np ? BucketOf(a.it.bk, Gather(a.it.cs, a.it.bk.start_v, p)) : kNoBucket;
      BucketAdd(a.it.bk, key, val);
    }
  }
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
  const uint64_t grid = GridFor(a.nchunks, kWavesPerBlock, max_grid);
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
