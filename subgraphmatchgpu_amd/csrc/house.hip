// house.hip -- house count from edge statistics (HouseArgs in kernels.h).
//
// Passes over a degree-ranked graph without self-loops:
//   1. triangles per edge: the clique row kernels with CliqueArgs::tc fill t
//      for every upper entry (an entry of some N+(x));
//   2. HouseMirrorKernel: one wave per vertex u gives each lower entry
//      (u -> v, v < u) the t of its mirror (v -> u), adds up u's wedge load
//      (the entries of each N(v) below u) and sums t(t-1) over u's upper
//      entries;
//   3. wedges: apexes with a load of at most kHouseHashMax are walked by one
//      wave each into an LDS hash table (HouseWedgeHashKernel); heavier ones
//      in groups, one dense u64 array over w per apex (HouseWedgeDenseKernel,
//      one wave per lower neighbour v, one atomic per wedge), reduced and
//      cleared by HouseWedgeReduceKernel.
//
// The reference counts the house by its LIGHT plan (match order 1, 2, 3, 0, 4;
// src/query/optimizer/LazyTraversalPlanHardCode.h:59-117), which enumerates
// every (v1, v2, v3) path; this pass reaches the same count from per-edge
// triangle counts and a degree-ordered 4-cycle walk.
#include "kernels.h"

namespace sme {

namespace {

constexpr int kThreads = 256;
constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;
constexpr uint64_t kMaxGrid = 16384;
constexpr uint32_t kHashSlots = 2 * kHouseHashMax;  // at most half full
constexpr int kHashBits = 10;
static_assert(kHashSlots == (1u << kHashBits), "hash index width");
constexpr uint32_t kEmpty = 0xffffffffu;
constexpr unsigned long long kSumMask = (1ull << kHouseShift) - 1;

using u64 = unsigned long long;

uint64_t Grid(uint64_t n, uint64_t per_block) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return g > kMaxGrid ? kMaxGrid : g;
}

__device__ __forceinline__ void Sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ u64 Sum64(u64 v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ void Flush(unsigned long long* acc, int j, u64 v, int lane) {
  v = Sum64(v);
  if (lane == 0 && v) atomicAdd(acc + j, v);
}

// First index in [lo, hi) with a[i] >= key.
__device__ __forceinline__ uint64_t Lower(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t key) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// (c - 1) * s of one packed accumulator.
__device__ __forceinline__ u64 Pairs(u64 p) {
  const u64 c = p >> kHouseShift;
  return c ? (c - 1) * (p & kSumMask) : 0;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void HouseMirrorKernel(HouseArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
  u64 sq = 0;
  for (uint64_t u = wave; u < a.n; u += nwaves) {
    const uint64_t b = a.rp[u], m = a.up[u], e = a.rp[u + 1];
    u64 load = 0;
    for (uint64_t k = b + lane; k < m; k += kWave) {
      const uint32_t v = a.cols[k];
      // u sits in N+(v); the entries of N(v) before it are v's neighbours below u
      const uint64_t pos = Lower(a.cols, a.up[v], a.rp[v + 1], static_cast<uint32_t>(u));
      a.t[k] = pos < a.rp[v + 1] && a.cols[pos] == u ? a.t[pos] : 0u;
      load += pos - a.rp[v];
    }
    for (uint64_t k = m + lane; k < e; k += kWave) {
      const u64 x = a.t[k];
      sq += x * (x ? x - 1 : 0);
    }
    load = Sum64(load);
    if (lane == 0) a.load[u] = load;
  }
  Flush(a.acc, 1, sq, lane);
}

__global__ __launch_bounds__(kThreads) void EdgePairsKernel(HouseArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
  u64 tri = 0, pairs = 0;
  for (uint64_t u = wave; u < a.n; u += nwaves) {
    for (uint64_t k = a.up[u] + lane; k < a.rp[u + 1]; k += kWave) {
      const u64 x = a.t[k];
      tri += x;
      pairs += x * (x ? x - 1 : 0) / 2;
    }
  }
  Flush(a.acc, 0, tri, lane);
  Flush(a.acc, 1, pairs, lane);
}

// One wave per apex u: every wedge u - v - w (v, w < u) goes into an LDS
// hash table keyed by w (linear probing, at most half full since u's load,
// and so its distinct w, is at most kHouseHashMax).
__global__ __launch_bounds__(kThreads) void HouseWedgeHashKernel(HouseArgs a) {
  __shared__ uint32_t lds_key[kWaves][kHashSlots];
  __shared__ u64 lds_val[kWaves][kHashSlots];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  uint32_t* key = lds_key[wid];
  u64* val = lds_val[wid];
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
  for (uint32_t j = lane; j < kHashSlots; j += kWave) {
    key[j] = kEmpty;
    val[j] = 0;
  }
  u64 sum = 0;
  for (uint64_t x = wave; x < a.napex; x += nwaves) {
    const uint32_t u = a.apex[x];
    Sync();
    for (uint64_t k = a.rp[u]; k < a.up[u]; ++k) {
      const uint32_t v = a.cols[k];
      const u64 tuv = a.t[k];
      const uint64_t end = Lower(a.cols, a.up[v], a.rp[v + 1], u);
      for (uint64_t e = a.rp[v] + lane; e < end; e += kWave) {
        const uint32_t w = a.cols[e];
        const u64 add = (1ull << kHouseShift) + tuv + a.t[e];
        uint32_t h = (w * 0x9E3779B1u) >> (32 - kHashBits);
        while (true) {
          const uint32_t old = atomicCAS(key + h, kEmpty, w);
          if (old == kEmpty || old == w) {
            atomicAdd(val + h, add);
            break;
          }
          h = (h + 1) & (kHashSlots - 1);
        }
      }
    }
    Sync();
    for (uint32_t j = lane; j < kHashSlots; j += kWave) {
      if (key[j] != kEmpty) {
        sum += Pairs(val[j]);
        key[j] = kEmpty;
        val[j] = 0;
      }
    }
  }
  Flush(a.acc, 0, sum, lane);
}

// Apex of dense item t: largest x with item_first[x] <= t.
__device__ __forceinline__ uint64_t ItemApex(const HouseArgs& a, uint64_t t) {
  uint64_t lo = 0, hi = a.napex - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (a.item_first[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// One wave per (apex u, lower neighbour v): N(v) below u, 64 entries per
// step, one packed atomic per wedge into u's dense array; for apexes reduced
// through a touched list, the w seen first (the atomic found 0) are appended.
__global__ __launch_bounds__(kThreads) void HouseWedgeDenseKernel(HouseArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const u64 below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
  for (uint64_t it = wave; it < a.nitems; it += nwaves) {
    const uint64_t x = ItemApex(a, it);
    const uint32_t u = a.apex[x];
    const uint64_t k = a.rp[u] + (it - a.item_first[x]);
    const uint32_t v = a.cols[k];
    const u64 tuv = a.t[k];
    const uint64_t end = Lower(a.cols, a.up[v], a.rp[v + 1], u);
    u64* slot = a.slots + x * a.n;
    const bool listed = x >= a.nfull;
    for (uint64_t base = a.rp[v]; base < end; base += kWave) {
      const uint64_t e = base + lane;
      const bool act = e < end;
      uint32_t w = 0;
      bool first = false;
      if (act) {
        w = a.cols[e];
        first = atomicAdd(slot + w, (1ull << kHouseShift) + tuv + a.t[e]) == 0;
      }
      if (listed) {
        const u64 m = __ballot(first);
        if (m) {
          u64 at = 0;
          if (lane == 0) at = atomicAdd(a.ntouched + x, static_cast<u64>(__popcll(m)));
          at = __shfl(at, 0);
          if (first) a.touched[x * a.n + at + __popcll(m & below)] = w;
        }
      }
    }
  }
}

// blockIdx.y = apex x: sums (c - 1) s over its w and clears them.
__global__ __launch_bounds__(kThreads) void HouseWedgeReduceKernel(HouseArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t x = blockIdx.y;
  u64* slot = a.slots + x * a.n;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  u64 sum = 0;
  if (x < a.nfull) {
    const uint64_t u = a.apex[x];
    for (uint64_t w = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; w < u; w += stride) {
      const u64 p = slot[w];
      if (p) {
        sum += Pairs(p);
        slot[w] = 0;
      }
    }
  } else {
    const uint64_t cnt = a.ntouched[x];
    const uint32_t* list = a.touched + x * a.n;
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < cnt; j += stride) {
      const uint32_t w = list[j];
      sum += Pairs(slot[w]);
      slot[w] = 0;
    }
  }
  Flush(a.acc, 0, sum, lane);
}

}  // namespace

void LaunchHouseMirror(const HouseArgs& a, hipStream_t stream) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(HouseMirrorKernel, dim3(Grid(a.n, kWaves)), dim3(kThreads), 0, stream, a);
}

void LaunchEdgePairs(const HouseArgs& a, hipStream_t stream) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(EdgePairsKernel, dim3(Grid(a.n, kWaves)), dim3(kThreads), 0, stream, a);
}

void LaunchHouseWedgeHash(const HouseArgs& a, hipStream_t stream) {
  if (a.napex == 0) return;
  hipLaunchKernelGGL(HouseWedgeHashKernel, dim3(Grid(a.napex, kWaves)), dim3(kThreads), 0, stream, a);
}

void LaunchHouseWedgeDense(const HouseArgs& a, hipStream_t stream) {
  if (a.nitems == 0) return;
  hipLaunchKernelGGL(HouseWedgeDenseKernel, dim3(Grid(a.nitems, kWaves)), dim3(kThreads), 0, stream, a);
}

void LaunchHouseWedgeReduce(const HouseArgs& a, hipStream_t stream) {
  if (a.napex == 0) return;
  hipLaunchKernelGGL(HouseWedgeReduceKernel, dim3(64, static_cast<uint32_t>(a.napex)), dim3(kThreads), 0,
                     stream, a);
}

}  // namespace sme
