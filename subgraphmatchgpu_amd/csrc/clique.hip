// clique.hip -- k-clique counting over bitmaps of upper neighbourhoods
// (CliqueArgs in kernels.h).
//
// Start vertices with at most 64 upper neighbours are counted by
// CliqueSmallKernel: one wave per vertex, lane i holding row i in a register,
// so every set below a row is one u64 and every step of the count is a
// readlane, an AND and a popcount.  Larger neighbourhoods are counted in
// batches: CliqueRowsKernel writes each vertex's rows to device memory (one
// wave per row, its lanes testing 64 candidates at a time by binary search
// in the row vertex's own upper list), then CliqueCountKernel takes one row
// i per wave, keeps the sets it derives from row i in LDS and counts the
// last level by AND + popcount over whole words, 64 clique members per word.
//
// The reference counts the same cliques by intersecting sorted adjacency
// lists level by level (src/gpu/impl/light/op/LIGHTCompute.cuh,
// LIGHTComputeCount.cuh:13-119); here the lists of one start vertex's
// neighbourhood become bit rows, and one word operation replaces up to 64
// membership probes.
#include "kernels.h"

namespace sme {

namespace {

constexpr int kThreads = 256;
constexpr int kWave = 64;
constexpr int kWaves = kThreads / kWave;
constexpr uint32_t kWordsMax = kLocalMax / 64;  // words of the longest row
static_assert(2 * kRowsHashSmall <= (1u << 10) && 2 * kRowsHashMax <= (1u << 13), "hash tables at most half full");
constexpr uint64_t kMaxGrid = 16384;

using u64 = unsigned long long;

uint64_t Grid(uint64_t n, uint64_t per_block) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return g > kMaxGrid ? kMaxGrid : g;
}

// Orders the LDS accesses of a wave: every lane's earlier accesses complete
// before any lane's later ones.
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

__device__ __forceinline__ uint32_t Lane32(uint32_t v, uint32_t lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
}

__device__ __forceinline__ u64 Lane64(u64 v, uint32_t lane) {
  return (static_cast<u64>(Lane32(static_cast<uint32_t>(v >> 32), lane)) << 32) |
         Lane32(static_cast<uint32_t>(v), lane);
}

__device__ __forceinline__ u64 First64(u64 v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<u64>(hi) << 32) | lo;
}

// Does the sorted a[lo, hi) hold key?  *pos: where (when it does).
__device__ __forceinline__ bool Holds(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t key,
                                      uint64_t* pos) {
  const uint64_t end = hi;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  *pos = lo;
  return lo < end && a[lo] == key;
}

// Largest bucket index b with bounds[b] <= v.
__device__ __forceinline__ uint64_t BucketIndex(const BucketRef& b, uint32_t v) {
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

__device__ __forceinline__ void AddBucket(const BucketRef& b, uint32_t v, u64 per_lane, int lane) {
  const u64 tot = Sum64(per_lane);
  if (lane == 0 && tot) atomicAdd(b.counts + BucketIndex(b, v), tot);
}

__device__ __forceinline__ void Flush(unsigned long long* acc, int j, u64 v, int lane) {
  v = Sum64(v);
  if (lane == 0 && v) atomicAdd(acc + j, v);
}

// ---------------------------------------------------------------------------
// One wave per start vertex with d <= 64.  Lane i builds row i: for each i
// the wave searches u_j (lane j > i) in N+(u_i) and the ballot is row i.
template <int K>
__global__ __launch_bounds__(kThreads) void CliqueSmallKernel(CliqueArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
  const uint64_t nv = a.v_end - a.v_begin;
  u64 c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0, c6 = 0;
  for (uint64_t t = wave; t < nv; t += nwaves) {
    const uint32_t v = a.v_begin + static_cast<uint32_t>(t);
    const uint64_t nb = a.up[v];
    const uint64_t d = a.rp[v + 1] - nb;
    if (d > kLocalSmall) {
      // (big == nullptr: the caller already holds these lists)
      if (lane == 0 && d <= kLocalMax) {
        if (a.big) {
          const u64 pos = atomicAdd(a.nbig, 1ull);
          a.big[pos] = v;
          a.big_d[pos] = static_cast<uint32_t>(d);
        }
        c1 += 1;
        c2 += d;
      } else if (lane == 0 && a.huge) {
        a.huge[atomicAdd(a.nhuge, 1ull)] = v;
      }
      continue;
    }
    if (lane == 0) {
      c1 += 1;
      c2 += d;
    }
    if (d < 2) continue;
    const uint32_t du = static_cast<uint32_t>(d);
    const uint32_t uj = static_cast<uint32_t>(lane) < du ? a.cols[nb + lane] : 0u;
    u64 row = 0;     // lane i: row i
    uint32_t col = 0;  // lane j: rows i < j holding bit j
    // rows i and i + 1 together: their two searches run in lockstep, so two
    // independent loads are in flight per step
    for (uint32_t i = 0; i + 1 < du; i += 2) {
      const uint32_t ua = Lane32(uj, i);
      const uint32_t ub = Lane32(uj, i + 1);
      uint64_t lo1 = a.up[ua], hi1 = a.rp[ua + 1], lo2 = a.up[ub], hi2 = a.rp[ub + 1];
      const uint64_t end1 = hi1, end2 = hi2;
      const bool act1 = static_cast<uint32_t>(lane) > i && static_cast<uint32_t>(lane) < du;
      const bool act2 = static_cast<uint32_t>(lane) > i + 1 && static_cast<uint32_t>(lane) < du;
      if (!act1) lo1 = hi1;
      if (!act2) lo2 = hi2;
      while (lo1 < hi1 || lo2 < hi2) {
        const bool g1 = lo1 < hi1, g2 = lo2 < hi2;
        const uint64_t m1 = (lo1 + hi1) >> 1, m2 = (lo2 + hi2) >> 1;
        const uint32_t x1 = g1 ? a.cols[m1] : 0u;
        const uint32_t x2 = g2 ? a.cols[m2] : 0u;
        if (g1) {
          if (x1 < uj)
            lo1 = m1 + 1;
          else
            hi1 = m1;
        }
        if (g2) {
          if (x2 < uj)
            lo2 = m2 + 1;
          else
            hi2 = m2;
        }
      }
      const bool ok1 = act1 && lo1 < end1 && a.cols[lo1] == uj;
      const bool ok2 = act2 && lo2 < end2 && a.cols[lo2] == uj;
      const u64 r1 = __ballot(ok1);
      const u64 r2 = __ballot(ok2);
      if (static_cast<uint32_t>(lane) == i) row = r1;
      if (static_cast<uint32_t>(lane) == i + 1) row = r2;
      if (a.tc) {
        // edges (u_i, u_j): triangles (v, u_i, u_j)
        if (ok1) {
          atomicAdd(a.tc + lo1, 1u);
          ++col;
        }
        if (ok2) {
          atomicAdd(a.tc + lo2, 1u);
          ++col;
        }
      }
    }
    const u64 tri = __popcll(row);
    // edge (v, u_i): the triangles (v, u_i, x) counted from v
    if (a.tc && static_cast<uint32_t>(lane) < du && tri + col)
      atomicAdd(a.tc + nb + lane, static_cast<uint32_t>(tri) + col);
    c3 += tri;
    u64 vk = K == 3 ? tri : 0;  // this lane's share of the k-cliques from v
    if (K >= 4) {
      for (uint32_t i = 0; i + 2 < du; ++i) {
        const u64 s1 = Lane64(row, i);
        if (!s1) continue;
        const u64 x4 = ((s1 >> lane) & 1) ? __popcll(s1 & row) : 0;
        c4 += x4;
        if (K == 4) {
          vk += x4;
          continue;
        }
        for (u64 m1 = s1; m1; m1 &= m1 - 1) {
          const u64 s2 = s1 & Lane64(row, static_cast<uint32_t>(__builtin_ctzll(m1)));
          if (!s2) continue;
          const u64 x5 = ((s2 >> lane) & 1) ? __popcll(s2 & row) : 0;
          c5 += x5;
          if (K == 5) {
            vk += x5;
            continue;
          }
          for (u64 m2 = s2; m2; m2 &= m2 - 1) {
            const u64 s3 = s2 & Lane64(row, static_cast<uint32_t>(__builtin_ctzll(m2)));
            if (!s3) continue;
            const u64 x6 = ((s3 >> lane) & 1) ? __popcll(s3 & row) : 0;
            c6 += x6;
            vk += x6;
          }
        }
      }
    }
    if (a.bk.bounds) AddBucket(a.bk, v, vk, lane);
  }
  Flush(a.acc, 1, c1, lane);
  Flush(a.acc, 2, c2, lane);
  Flush(a.acc, 3, c3, lane);
  if (K >= 4) Flush(a.acc, 4, c4, lane);
  if (K >= 5) Flush(a.acc, 5, c5, lane);
  if (K >= 6) Flush(a.acc, 6, c6, lane);
}

// ---------------------------------------------------------------------------
// Big vertices: row items.  Vertex x of the batch owns items
// [row_first[x], row_first[x+1]) - row_first[0]; the wave-uniform search
// finds it.
__device__ __forceinline__ uint64_t ItemVertex(const CliqueArgs& a, uint64_t t) {
  const uint64_t base = a.row_first[0];
  uint64_t lo = 0, hi = a.nverts - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (a.row_first[mid] - base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// One wave per row: word c of row i is the ballot of "u_j in N+(u_i)" over
// j = 64c + lane (j > i).  Words are gathered one per lane and stored 64 at
// a time.
__global__ __launch_bounds__(kThreads) void CliqueRowsKernel(CliqueArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
  u64 c3 = 0;  // triangles, when no count kernel follows (per-edge counts)
  for (uint64_t t = wave; t < a.nitems; t += nwaves) {
    const uint64_t x = ItemVertex(a, t);
    const uint32_t v = a.verts[x];
    const uint32_t d = static_cast<uint32_t>(a.row_first[x + 1] - a.row_first[x]);
    const uint32_t i = static_cast<uint32_t>(t - (a.row_first[x] - a.row_first[0]));
    const uint32_t W = (d + 63) / 64;
    u64* rows = a.rows + (a.word_first[x] - a.word_first[0]);  // word w of row i at rows[w * d + i]
    const uint64_t nb = a.up[v];
    const uint32_t ui = a.cols[nb + i];
    const uint64_t lb = a.up[ui], le = a.rp[ui + 1];
    u64 keep = 0;
    uint32_t bits = 0;
    for (uint32_t c = 0; c < W; ++c) {
      const uint32_t j = c * 64 + static_cast<uint32_t>(lane);
      uint64_t pos = 0;
      const bool ok = j > i && j < d && Holds(a.cols, lb, le, a.cols[nb + j], &pos);
      const u64 w = __ballot(ok);
      if ((c & 63u) == static_cast<uint32_t>(lane)) keep = w;
      if ((c & 63u) == 63u || c + 1 == W) {
        if (static_cast<uint32_t>(lane) <= (c & 63u))
          rows[static_cast<uint64_t>((c & ~63u) + lane) * d + i] = keep;
      }
      if (ok && a.tc) {
        atomicAdd(a.tc + pos, 1u);     // edge (u_i, u_j)
        atomicAdd(a.tc + nb + j, 1u);  // edge (v, u_j)
      }
      bits += static_cast<uint32_t>(__popcll(w));
    }
    // edge (v, u_i): one triangle per bit of row i
    if (a.tc && lane == 0 && bits) atomicAdd(a.tc + nb + i, bits);
    if (lane == 0) c3 += bits;
  }
  if (a.tc) Flush(a.acc, 3, c3, lane);
}

// Chunk t of a batch: largest x with chunk_first[x] - chunk_first[0] <= t.
__device__ __forceinline__ uint64_t ChunkVertex(const CliqueArgs& a, uint64_t t) {
  const uint64_t base = a.chunk_first[0];
  uint64_t lo = 0, hi = a.nverts - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (a.chunk_first[mid] - base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// One workgroup per chunk of 64 rows of one vertex v.  N+(v) goes into an LDS
// hash table (value -> local index j, at most half full); each wave then
// builds rows of the chunk by streaming N+(u_i) up to the largest value of
// N+(v), 64 consecutive entries per load, and setting bit j of the row (an
// LDS OR) for every entry found.  Every entry of N+(u_i) lies above u_i, so
// the bits set are exactly those j > i with u_j in N+(u_i).
template <int kBits>
__global__ __launch_bounds__(kThreads) void CliqueRowsHashKernel(CliqueArgs a) {
  constexpr uint32_t kSlots = 1u << kBits;
  constexpr uint32_t kEmptyKey = 0xffffffffu;
  __shared__ uint32_t lds_key[kSlots];
  __shared__ uint16_t lds_idx[kSlots];
  __shared__ u64 lds_row[kWaves][kRowsHashMax / 64];
  // per-edge counts (tc): the chunk's triangles through edge (v, u_j),
  // flushed with one atomic per j
  __shared__ uint32_t lds_col[kSlots / 2];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  u64* row = lds_row[wid];
  u64 c3 = 0;
  for (uint64_t t = blockIdx.x; t < a.nchunks; t += gridDim.x) {
    const uint64_t x = ChunkVertex(a, t);
    const uint32_t v = a.verts[x];
    const uint32_t d = static_cast<uint32_t>(a.row_first[x + 1] - a.row_first[x]);
    const uint32_t W = (d + 63) / 64;
    const uint32_t r0 = static_cast<uint32_t>(t - (a.chunk_first[x] - a.chunk_first[0])) * 64;
    const uint32_t r1 = r0 + 64 < d ? r0 + 64 : d;
    const uint64_t nb = a.up[v];
    u64* rows = a.rows + (a.word_first[x] - a.word_first[0]);
    __syncthreads();  // the previous chunk's lookups are done
    for (uint32_t k = threadIdx.x; k < kSlots; k += kThreads) lds_key[k] = kEmptyKey;
    if (a.tc)
      for (uint32_t k = threadIdx.x; k < d; k += kThreads) lds_col[k] = 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < d; j += kThreads) {
      const uint32_t w = a.cols[nb + j];
      uint32_t h = (w * 0x9E3779B1u) >> (32 - kBits);
      while (atomicCAS(lds_key + h, kEmptyKey, w) != kEmptyKey) h = (h + 1) & (kSlots - 1);
      lds_idx[h] = static_cast<uint16_t>(j);
    }
    __syncthreads();
    const uint32_t vmax = a.cols[nb + d - 1];
    // the next row's vertex and list bounds are loaded while this row streams
    uint32_t i = r0 + wid;
    uint64_t lb_next = 0, le_next = 0;
    if (i < r1) {
      const uint32_t u = a.cols[nb + i];
      lb_next = a.up[u];
      le_next = a.rp[u + 1];
    }
    for (; i < r1; i += kWaves) {
      const uint64_t lb = lb_next, le = le_next;
      if (i + kWaves < r1) {
        const uint32_t u = a.cols[nb + i + kWaves];
        lb_next = a.up[u];
        le_next = a.rp[u + 1];
      }
      for (uint32_t w = lane; w < W; w += kWave) row[w] = 0;
      Sync();
      uint32_t bits = 0;
      // chunks of 64 entries, the next chunk's load issued before this
      // chunk's lookups
      uint64_t e0 = lb;
      uint32_t wcur = lb + lane < le ? a.cols[lb + lane] : kEmptyKey;
      while (e0 < le) {
        // N+(u_i) is sorted: nothing from a chunk starting above the
        // largest value of N+(v) on can be found (lane 0 holds the chunk's
        // first entry)
        if (static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(wcur))) > vmax) break;
        const uint64_t n0 = e0 + kWave;
        const uint32_t wnext = n0 + lane < le ? a.cols[n0 + lane] : kEmptyKey;
        const uint64_t e = e0 + lane;
        const uint32_t w = wcur;
        e0 = n0;
        wcur = wnext;
        if (e < le) {
          uint32_t h = (w * 0x9E3779B1u) >> (32 - kBits);
          uint32_t k;
          while ((k = lds_key[h]) != w && k != kEmptyKey) h = (h + 1) & (kSlots - 1);
          if (k == w) {
            const uint32_t j = lds_idx[h];
            atomicOr(row + (j >> 6), 1ull << (j & 63));
            ++bits;
            if (a.tc) {
              atomicAdd(a.tc + e, 1u);     // edge (u_i, u_j)
              atomicAdd(lds_col + j, 1u);  // edge (v, u_j)
            }
          }
        }
      }
      Sync();
      for (uint32_t w = lane; w < W; w += kWave) rows[static_cast<uint64_t>(w) * d + i] = row[w];
      if (a.tc) {
        // edge (v, u_i): one triangle per bit of row i (only the per-edge
        // pass needs the row's popcount here; otherwise the count kernel
        // takes it)
        const u64 tot = Sum64(bits);
        if (lane == 0) {
          if (tot) atomicAdd(a.tc + nb + i, static_cast<uint32_t>(tot));
          c3 += tot;
        }
      }
      Sync();
    }
    if (a.tc) {
      __syncthreads();
      for (uint32_t j = threadIdx.x; j < d; j += kThreads)
        if (lds_col[j]) atomicAdd(a.tc + nb + j, lds_col[j]);
    }
  }
  if (a.tc) Flush(a.acc, 3, c3, lane);
}

// |s & row_j| over the words of row j that can be non-zero (bits above j).
// Rows are stored word-major (word w of row j at rows[w * d + j]), so the
// lanes of a probe round, holding nearby members j, read nearby addresses.
__device__ __forceinline__ u64 Probe(const u64* s, uint32_t W, uint32_t d, const u64* rows, uint32_t j) {
  const u64* r = rows + j;
  u64 x = 0;
  for (uint32_t w = (j + 1) >> 6; w < W; ++w) x += __popcll(s[w] & r[static_cast<uint64_t>(w) * d]);
  return x;
}

// Sum over the members j of s (words [lo, W), in LDS) of |s & row_j|, as
// per-lane partial sums.  Members are queued in LDS (q: 2 x 64 entries) so
// every probe round has 64 of them, whatever the density of s.
__device__ __forceinline__ u64 LastLevel(const u64* s, uint32_t lo, uint32_t W, uint32_t d, const u64* rows, uint32_t* q,
                         int lane) {
  const u64 below = lane ? (~0ull >> (64 - lane)) : 0ull;
  u64 sum = 0;
  uint32_t qn = 0;
  for (uint32_t c = lo; c < W; ++c) {
    const u64 bits = First64(s[c]);
    if (!bits) continue;
    if ((bits >> lane) & 1) q[qn + __popcll(bits & below)] = c * 64 + static_cast<uint32_t>(lane);
    qn += __popcll(bits);
    if (qn >= static_cast<uint32_t>(kWave)) {
      Sync();
      sum += Probe(s, W, d, rows, q[lane]);
      const uint32_t rest = qn - kWave;
      const uint32_t carry = static_cast<uint32_t>(lane) < rest ? q[kWave + lane] : 0u;
      Sync();
      if (static_cast<uint32_t>(lane) < rest) q[lane] = carry;
      qn = rest;
      Sync();
    }
  }
  if (qn) {
    Sync();
    if (static_cast<uint32_t>(lane) < qn) sum += Probe(s, W, d, rows, q[lane]);
  }
  Sync();
  return sum;
}

struct CliqueSums {
  u64 c3 = 0, c4 = 0, c5 = 0, c6 = 0;
};

// The cliques from row i of a vertex with d upper neighbours whose rows are
// stored word-major at `rows` (device memory or LDS): s1 = row i, and for
// K >= 5 every s2 = s1 & row j (j in s1), s3 = s2 & row l, each kept in the
// wave's LDS buffers while the members below it are walked; the last set's
// members are probed by LastLevel.  Adds the level counts to *cs and
// returns this lane's share of the k-cliques.
template <int K>
__device__ __forceinline__ u64 CountRow(const u64* rows, uint32_t d, uint32_t i, u64* s1, u64* s2, u64* s3,
                                        uint32_t* q, int lane, CliqueSums* cs) {
  const uint32_t W = (d + 63) / 64;
  const u64* ri = rows + i;
  const uint32_t lo1 = (i + 1) >> 6;
  u64 vk = 0;
  Sync();
  for (uint32_t w = lo1 + lane; w < W; w += kWave) {
    const u64 xw = ri[static_cast<uint64_t>(w) * d];
    if (K >= 4) s1[w] = xw;
    cs->c3 += __popcll(xw);
    if (K == 3) vk += __popcll(xw);
  }
  if (K < 4) return vk;
  Sync();
  if (K == 4) {
    const u64 y = LastLevel(s1, lo1, W, d, rows, q, lane);
    cs->c4 += y;
    return vk + y;
  }
  for (uint32_t c = lo1; c < W; ++c) {
    for (u64 m1 = First64(s1[c]); m1; m1 &= m1 - 1) {
      const uint32_t j = c * 64 + static_cast<uint32_t>(__builtin_ctzll(m1));
      const u64* rj = rows + j;
      const uint32_t lo2 = (j + 1) >> 6;
      u64 nz = 0;
      Sync();
      for (uint32_t w = lo2 + lane; w < W; w += kWave) {
        const u64 xw = s1[w] & rj[static_cast<uint64_t>(w) * d];
        s2[w] = xw;
        cs->c4 += __popcll(xw);
        nz |= xw;
      }
      if (!__ballot(nz != 0)) continue;
      Sync();
      if (K == 5) {
        const u64 y = LastLevel(s2, lo2, W, d, rows, q, lane);
        cs->c5 += y;
        vk += y;
        continue;
      }
      for (uint32_t c2 = lo2; c2 < W; ++c2) {
        for (u64 m2 = First64(s2[c2]); m2; m2 &= m2 - 1) {
          const uint32_t l = c2 * 64 + static_cast<uint32_t>(__builtin_ctzll(m2));
          const u64* rl = rows + l;
          const uint32_t lo3 = (l + 1) >> 6;
          u64 nz3 = 0;
          Sync();
          for (uint32_t w = lo3 + lane; w < W; w += kWave) {
            const u64 xw = s2[w] & rl[static_cast<uint64_t>(w) * d];
            s3[w] = xw;
            cs->c5 += __popcll(xw);
            nz3 |= xw;
          }
          if (!__ballot(nz3 != 0)) continue;
          Sync();
          const u64 y = LastLevel(s3, lo3, W, d, rows, q, lane);
          cs->c6 += y;
          vk += y;
        }
      }
    }
  }
  return vk;
}

__device__ __forceinline__ void FlushSums(unsigned long long* acc, int K, CliqueSums* cs, int lane) {
  Flush(acc, 3, cs->c3, lane);
  if (K >= 4) Flush(acc, 4, cs->c4, lane);
  if (K >= 5) Flush(acc, 5, cs->c5, lane);
  if (K >= 6) Flush(acc, 6, cs->c6, lane);
}

// One wave per row i of a big vertex, rows in device memory.
template <int K>
__global__ __launch_bounds__(kThreads) void CliqueCountKernel(CliqueArgs a) {
  constexpr int kBufs = K >= 4 ? K - 3 : 1;
  __shared__ u64 lds_s[kWaves][kBufs][kWordsMax];
  __shared__ uint32_t lds_q[kWaves][2 * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  u64* s1 = lds_s[wid][0];
  u64* s2 = lds_s[wid][kBufs > 1 ? 1 : 0];
  u64* s3 = lds_s[wid][kBufs > 2 ? 2 : 0];
  uint32_t* q = lds_q[wid];
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * blockDim.x / kWave;
  CliqueSums cs;
  // up to 4-cliques a row is light next to the search for its vertex: a
  // wave takes groups of kRowGroup consecutive rows, one search for the
  // group's first vertex, then a walk forward.  Deeper cliques' rows are
  // heavy, and the first rows of the largest vertices heaviest; one row at
  // a time keeps them spread over the waves.
  constexpr uint64_t kRowGroup = K <= 4 ? 16 : 1;
  const uint64_t base = a.row_first[0];
  for (uint64_t t0 = wave * kRowGroup; t0 < a.nitems; t0 += nwaves * kRowGroup) {
    const uint64_t te = t0 + kRowGroup < a.nitems ? t0 + kRowGroup : a.nitems;
    uint64_t x = ItemVertex(a, t0);
    for (uint64_t t = t0; t < te; ++t) {
      while (a.row_first[x + 1] - base <= t) ++x;
      const uint32_t d = static_cast<uint32_t>(a.row_first[x + 1] - a.row_first[x]);
      const uint32_t i = static_cast<uint32_t>(t - (a.row_first[x] - base));
      const u64* rows = a.rows + (a.word_first[x] - a.word_first[0]);
      const u64 vk = CountRow<K>(rows, d, i, s1, s2, s3, q, lane, &cs);
      if (a.bk.bounds) AddBucket(a.bk, a.verts[x], vk, lane);
    }
  }
  FlushSums(a.acc, K, &cs, lane);
}

}  // namespace

void LaunchCliqueSmall(const CliqueArgs& a, hipStream_t stream) {
  if (a.v_end <= a.v_begin) return;
  const dim3 grid(Grid(a.v_end - a.v_begin, kWaves)), block(kThreads);
  switch (a.k) {
    case 3: hipLaunchKernelGGL(CliqueSmallKernel<3>, grid, block, 0, stream, a); break;
    case 4: hipLaunchKernelGGL(CliqueSmallKernel<4>, grid, block, 0, stream, a); break;
    case 5: hipLaunchKernelGGL(CliqueSmallKernel<5>, grid, block, 0, stream, a); break;
    case 6: hipLaunchKernelGGL(CliqueSmallKernel<6>, grid, block, 0, stream, a); break;
  }
}

void LaunchCliqueRows(const CliqueArgs& a, hipStream_t stream) {
  if (a.nitems == 0) return;
  hipLaunchKernelGGL(CliqueRowsKernel, dim3(Grid(a.nitems, kWaves)), dim3(kThreads), 0, stream, a);
}

void LaunchCliqueRowsHash(const CliqueArgs& a, bool small_table, hipStream_t stream) {
  if (a.nchunks == 0) return;
  const dim3 grid(Grid(a.nchunks, 1)), block(kThreads);
  if (small_table)
    hipLaunchKernelGGL(CliqueRowsHashKernel<10>, grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL(CliqueRowsHashKernel<13>, grid, block, 0, stream, a);
}

void LaunchCliqueCount(const CliqueArgs& a, hipStream_t stream) {
  if (a.nitems == 0) return;
  const dim3 grid(Grid(a.nitems, kWaves)), block(kThreads);
  switch (a.k) {
    case 3: hipLaunchKernelGGL(CliqueCountKernel<3>, grid, block, 0, stream, a); break;
    case 4: hipLaunchKernelGGL(CliqueCountKernel<4>, grid, block, 0, stream, a); break;
    case 5: hipLaunchKernelGGL(CliqueCountKernel<5>, grid, block, 0, stream, a); break;
    case 6: hipLaunchKernelGGL(CliqueCountKernel<6>, grid, block, 0, stream, a); break;
  }
}

}  // namespace sme
