// kernels.h -- argument blocks and host launchers of the gfx950 kernels.
//
// Every kernel takes its pointers and plan data as one by-value argument
// block (kernarg segment), so no per-op pointer tables are copied to the
// device (the reference uploads them before every op,
// src/gpu/impl/light/LIGHTCommon.cuh:134-144).
//
// Work layout used by the item kernels: a level is a set of segments
// (partial instances, "paths"); segment s owns items [offs[s], offs[s+1]).
// One wave64 handles one chunk of 64 consecutive items; chunk_seg[c] is the
// segment of the chunk's first item, so a lane finds its segment with a short
// binary search over offs restricted to [chunk_seg[c], chunk_seg[c+1]].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sme {

constexpr int kQ = 6;  // kMaxQueryVerticesNum

// Materialized columns of a frame.  inst[w] == nullptr means the column is
// the implicit sequence iota[w] + path (the first matched vertex).
struct ColSet {
  const uint32_t* inst[kQ];
  uint32_t iota[kQ];
};

// Candidate set of one query vertex: segment k holds base[beg[k] ..
// beg[k]+len[k]); path p uses segment idx[p] (idx == nullptr: segment p).
struct CandRef {
  const uint32_t* base;
  const uint64_t* beg;
  const uint64_t* len;
  const uint32_t* idx;
};

// Per-path set-up of an intersection (COMPUTE / COMPUTE_PATH_COUNT /
// COMPUTE_COUNT).  The pivot is the backward list of smallest degree; it is
// cut to the value window implied by the < / > conditions.
//
// Intersection reuse: when `has_reuse` is set, the parent's candidate segment
// `reuse` (already the intersection of some of this vertex's backward lists)
// is one more list; `bconn` then holds only the lists it does not cover.  The
// pivot id `nb` denotes the reused segment.
struct PrepComputeArgs {
  uint64_t npaths;
  const uint64_t* rp;
  const uint32_t* cols;
  ColSet cs;
  CandRef reuse;
  uint8_t has_reuse;
  uint8_t nb;
  uint8_t bconn[kQ];
  uint8_t nlt, ngt;
  uint8_t lt_v[kQ];  // candidate < M[lt_v]
  uint8_t gt_v[kQ];  // candidate > M[gt_v]
  uint64_t* size;    // out: items of the path (restricted pivot length)
  uint64_t* pstart;  // out: first item's index into cols
  uint8_t* piv;      // out: which backward list is the pivot
  // Streaming pivot: with a reused segment of at most kStageEntries values
  // and one adjacency list of length <= stream_ratio x that segment, the
  // adjacency list becomes the pivot (read in order) and the short segment
  // is what gets probed (staged in LDS by the verify fast path); 0: the
  // shortest list is the pivot.
  uint32_t stream_ratio;
  // paths whose reused segment holds at most group_cap values are counted
  // by GroupCountKernel: they get no items here (0: none)
  uint32_t group_cap;
  // paths with at least wg_min pivot items (0: none) get no items either:
  // (path, items) is appended to wg_list and PathWorkgroupKernel counts each
  // with a whole workgroup; wg_count receives the number appended
  uint64_t wg_min;
  uint64_t* wg_list;
  unsigned long long* wg_count;
};

constexpr uint32_t kWgStage = 8192;  // LDS entries a workgroup stages for one path

constexpr uint32_t kStageEntries = 1024;  // LDS entries per wave for a staged list

// Per-path set-up of a scan over an existing candidate set (MATERIALIZE,
// FILTER_COMPUTE): the segment is cut to the < / > window.
struct PrepListArgs {
  uint64_t npaths;
  ColSet cs;
  CandRef cand;
  uint8_t nlt, ngt;
  uint8_t lt_v[kQ];
  uint8_t gt_v[kQ];
  uint64_t* size;
  uint64_t* pstart;
};

// Per-bucket attribution of counts: bucket k holds the paths whose first
// matched vertex (column `start_v`) lies in [bounds[k], bounds[k+1]).  Paths
// are ordered by start vertex at every level, so the lanes of a wave hold
// non-decreasing bucket ids and one segmented wave sum per run suffices.
struct BucketRef {
  const uint32_t* bounds;  // nullptr: no per-bucket counts
  uint64_t nbuckets;
  unsigned long long* counts;
  uint8_t start_v;
  // bucket of start vertex v is table[v - bounds[0]] (one load instead of
  // a binary search over bounds); nullptr: search bounds
  const uint32_t* table;
  uint32_t base;
};

// COMPUTE_COUNT of paths that extend a reused parent segment S with one
// adjacency list N(M[adj_v]) (the last level of a clique, for instance).
// Consecutive paths with the same segment -- the children of one parent --
// form a group.  One wave stages S in LDS once per group (sorted, plus a
// hash table of it), then streams every child's window of
// N(M[adj_v]) in order, 64 consecutive entries per load, and counts the
// values found in S.  Reads of the graph are coalesced instead of one
// random line per membership probe.  Paths whose segment holds more than
// `cap` values are left to the item schedule.
struct GroupCountArgs {
  uint64_t npaths;
  ColSet cs;
  CandRef seg;
  uint8_t adj_v;
  const uint64_t* rp;
  const uint32_t* cols;
  const uint64_t* up;  // first neighbour above v (nullptr: binary search)
  uint8_t nlt, ngt;
  uint8_t lt_v[kQ];
  uint8_t gt_v[kQ];
  uint8_t nne;
  uint8_t ne_v[2];
  uint32_t cap;
  unsigned long long* acc;
  unsigned long long* acc_level;
  // profiling (nullable): values streamed, and the sum over paths of
  // min(|S in window|, |list window|) (the walk of a pivot-and-probe
  // intersection, used for algorithmic bytes)
  unsigned long long* walked;
  unsigned long long* lower;
  BucketRef bk;
};
constexpr uint32_t kGroupMax = 512;  // longest segment a wave stages

// Item kernels.  `offs` / per-segment arrays are already offset to the
// batch's first segment; items are numbered from offs[0].
struct ItemArgs {
  uint64_t nitems;
  uint64_t nseg;
  uint64_t nchunks;
  const uint64_t* offs;       // nseg + 1
  const uint32_t* chunk_seg;  // nchunks
  // value of an item: src[pstart[s] + rank]
  const uint32_t* src;
  const uint64_t* pstart;
  // membership tests (intersection): every backward list but the pivot
  const uint64_t* rp;
  const uint32_t* cols;
  const uint8_t* piv;
  uint8_t nb;
  uint8_t bconn[kQ];
  // reused parent segment (see PrepComputeArgs); piv[s] == nb selects it
  CandRef reuse;
  uint8_t has_reuse;
  // adjacency bitmaps of high-degree vertices: membership of x in N(v) is
  // one bit when hub_id[v] != kNoHub (hub_id == nullptr: none built)
  const uint32_t* hub_id;
  const uint32_t* hub_bits;
  uint64_t hub_words;
  // non-equality conditions that the window cannot express
  uint8_t nne;
  uint8_t ne_v[kQ];
  ColSet cs;
  // COMPUTE_COUNT with one other unmaterialized vertex
  CandRef pair;
  uint8_t pair_op;
  // values the pair's set excludes without having been copied: bit k of
  // pair_excl_in[s] says M[pair_excl_v[k]] lies in segment s and is to be
  // left out (a FILTER_COMPUTE of != conditions kept as a view)
  uint8_t pair_nexcl;
  uint8_t pair_excl_v[2];
  const uint8_t* pair_excl_in;
  // COUNT over two unmaterialized vertices: `pair` is the first vertex's
  // set, `second` the other; swapped[s] == 1 when the second is smaller.
  CandRef second;
  const uint8_t* swapped;
  // outputs
  unsigned long long* mask;       // per chunk validity bits
  uint32_t* ccount;               // per chunk popcount
  unsigned long long* acc;        // total count
  unsigned long long* acc_level;  // this level's contribution
  BucketRef bk;                   // counting modes only
  // verify kernels: a wave takes super_chunks consecutive chunks; when they
  // all lie in one segment the per-segment data is loaded once and list
  // windows are staged in LDS (0 or 1: one chunk at a time)
  uint32_t super_chunks;
};

// COMPUTE_COUNT set-up with lane-per-path counting of short paths: paths
// whose pivot window holds at most lane_max items are verified in the set-up
// kernel by one lane each (their size becomes 0); `it` carries the membership
// data indexed by path (cs / reuse / pair of the whole frame).
struct PrepCountArgs {
  PrepComputeArgs pa;
  ItemArgs it;
  uint64_t lane_max;
  uint8_t pair;                  // COMPUTE_COUNT with one other unmaterialized vertex
  unsigned long long* lane_items;  // out (optional): items verified by lanes
};

enum VerifyMode : int { kModeMask = 0, kModeSum = 1, kModeSumPair = 2, kModeSumTwo = 3 };

// Rank structures built from a verify pass.
struct RankRef {
  const unsigned long long* mask;   // nullptr: every item survives
  const unsigned long long* cbase;  // exclusive scan of ccount (nchunks + 1)
  uint64_t nchunks;
  uint64_t nitems;
};

// MATERIALIZE output columns.
struct ExpandArgs {
  ItemArgs it;
  RankRef rk;
  uint32_t* out_new;              // column of the materialized vertex
  uint8_t ncarry;
  uint8_t carry_v[kQ];            // materialized columns copied to children
  uint32_t* carry_out[kQ];
  uint8_t nidx;
  CandRef idx_in[kQ];             // candidate indices copied to children
  uint32_t* idx_out[kQ];
};

// COMPUTE / FILTER_COMPUTE output: compacted values + per-segment window.
struct CompactArgs {
  ItemArgs it;
  RankRef rk;
  uint32_t* out_vals;
  uint64_t* out_beg;
  uint64_t* out_len;
};

// Per-path closed-form / single-set counts (COUNT).
struct CountPathArgs {
  uint64_t npaths;
  uint8_t mode;  // 0: direct closed form, 1: size of one set, 2: one per path
  int query_type;
  uint8_t ncun;
  uint8_t cun[kQ];
  CandRef cand[kQ];  // indexed by query vertex
  unsigned long long* acc;
  unsigned long long* acc_level;
  ColSet cs;         // start-vertex column for bucket attribution
  BucketRef bk;
};

// FILTER_COMPUTE of at most two != conditions kept as a view: per path,
// bit k of excl_in = [M[ne_v[k]] lies in the path's segment]; the level's
// output (segment size minus the excluded values present) goes to acc_level.
struct ExclusionArgs {
  uint64_t npaths;
  ColSet cs;
  CandRef cand;
  uint8_t nne;
  uint8_t ne_v[2];
  uint8_t* excl_in;
  unsigned long long* acc_level;
};

// Per-path set-up of the two-set COUNT.
struct PrepTwoArgs {
  uint64_t npaths;
  CandRef a, b;
  uint64_t* size;
  uint64_t* pstart;
  uint8_t* swapped;
};

// ---------------------------------------------------------------------------
// Launchers (all on `stream`).
void LaunchPrepCompute(const PrepComputeArgs& a, hipStream_t stream);
void LaunchPrepCount(const PrepCountArgs& a, hipStream_t stream);
void LaunchPrepList(const PrepListArgs& a, hipStream_t stream);
void LaunchPrepTwo(const PrepTwoArgs& a, hipStream_t stream);
void LaunchExclusion(const ExclusionArgs& a, hipStream_t stream);
void LaunchCountPaths(const CountPathArgs& a, hipStream_t stream);
void LaunchChunkPartition(const uint64_t* offs, uint64_t nseg, uint64_t nchunks,
                          uint32_t* chunk_seg, hipStream_t stream);
// max_grid caps the workgroups of the launch (0: default cap); every wave
// loops over chunks, so any cap >= 1 covers all items.
void LaunchVerify(const ItemArgs& a, int mode, uint32_t max_grid, hipStream_t stream);
void LaunchCompact(const CompactArgs& a, hipStream_t stream);
void LaunchGroupCount(const GroupCountArgs& a, uint32_t max_grid, hipStream_t stream);
// Counting level (mode kModeSum or kModeSumPair) of the paths listed by the
// set-up kernel (wg_list / wg_count, read on the device): one workgroup per
// path stages a window of one list in LDS and walks the pivot with 256
// threads.  items (nullable) receives the pivot items walked.
void LaunchPathWorkgroup(const ItemArgs& a, int mode, const uint64_t* wg_list,
                         const unsigned long long* wg_count, unsigned long long* items,
                         uint32_t max_grid, hipStream_t stream);
// up[v] = first index of rp[v]..rp[v+1] holding a neighbour > v.
void LaunchUpperStart(const uint64_t* rp, const uint32_t* cols, uint64_t n, uint64_t* up,
                      hipStream_t stream);
void LaunchExpand(const ExpandArgs& a, hipStream_t stream);
constexpr uint32_t kNoHub = 0xffffffffu;
// Sets bit x of row h for every x in N(hubs[h]); rows must be zero.
void LaunchBuildHubBits(const uint64_t* rp, const uint32_t* cols, const uint32_t* hubs,
                        uint64_t nhub, uint64_t words, uint32_t* bits, hipStream_t stream);
// out[p * nv + w] = column w of path p (nv columns, all materialized).
void LaunchGatherTuples(const ColSet& cs, int nv, uint64_t n, uint32_t* out, hipStream_t stream);
void LaunchFill32(uint32_t* out, uint64_t n, uint32_t value, hipStream_t stream);
// table[v - bounds[0]] = bucket of v for v in [bounds[0], bounds[nbuckets]).
void LaunchBucketTable(const uint32_t* bounds, uint64_t nbuckets, uint64_t range, uint32_t* table,
                       hipStream_t stream);
void LaunchIota32(uint32_t* out, uint64_t n, uint32_t base, hipStream_t stream);
void LaunchFill64(uint64_t* out, uint64_t n, uint64_t value, hipStream_t stream);
// Greedy batch split over paths [start, npaths): writes up to max_out pairs
// (right end, offs[right end]) to `bounds` and the pair count to counters[0].
void LaunchSplitBatches(const uint64_t* offs, uint64_t start, uint64_t npaths,
                        uint64_t bytes_per_item, uint64_t bytes_per_path, uint64_t budget,
                        uint64_t max_items, uint64_t max_paths, uint64_t* bounds,
                        uint64_t max_out, uint64_t* counters, hipStream_t stream);

// ---------------------------------------------------------------------------
// k-clique counting over bitmaps of upper neighbourhoods (clique.hip).
//
// With ordering, a k-clique plan counts each clique once, from its lowest
// vertex v, and every other vertex of the clique lies in N+(v) =
// cols[up[v], rp[v+1]) (the neighbours above v, sorted).  The count from v
// is therefore the number of (k-1)-cliques of the graph induced on N+(v).
// With d = |N+(v)| and N+(v) = {u_0 < ... < u_{d-1}}, that graph is held as
// d rows of W = ceil(d/64) u64 words: bit j of row i is set when j > i and
// u_j is a neighbour of u_i.  Cliques are then counted by AND + popcount:
//   triangles from v   = sum_i |row_i|
//   4-cliques from v   = sum_i sum_{j in row_i} |row_i & row_j|
//   5-cliques from v   = sum over j in row_i, l in row_i & row_j of
//                        |row_i & row_j & row_l|, and so on.
// The j-cliques (j = 1..k) whose lowest vertex is counted are exactly the
// per-level outputs of the clique plan (its level L outputs j-cliques with
// j = matched vertices after L), so acc[j] collects them.
constexpr uint32_t kLocalSmall = 64;    // d <= 64: one wave per start vertex, one word per row
constexpr uint32_t kLocalMax = 16384;   // d > kLocalMax: left to the level engine
constexpr int kCliqueMaxK = 6;

struct CliqueArgs {
  const uint64_t* rp;
  const uint32_t* cols;
  const uint64_t* up;
  int k;                         // clique size, 3..kCliqueMaxK
  unsigned long long* acc;       // acc[j] += j-cliques counted, j = 1..k (acc[0] unused)
  // per-start-vertex buckets (bounds == nullptr: none); counts[b] += the
  // k-cliques whose lowest vertex lies in bucket b
  BucketRef bk;
  // CliqueSmallKernel: start vertices [v_begin, v_end).  It counts those
  // with d <= kLocalSmall and adds acc[1] and acc[2] (1 and d) for every
  // vertex with d <= kLocalMax; vertices with kLocalSmall < d <= kLocalMax
  // are appended to (big, big_d), those with d > kLocalMax to huge (not when
  // big / huge are null: the caller holds these lists already).
  uint32_t v_begin, v_end;
  uint32_t* big;
  uint32_t* big_d;
  unsigned long long* nbig;
  uint32_t* huge;
  unsigned long long* nhuge;
  // CliqueRowsKernel / CliqueCountKernel: one batch of big vertices.
  // Vertex t's rows are row items [row_first[t], row_first[t+1]) (one per
  // row) and words [word_first[t], word_first[t+1]) - word_first[0] of rows.
  const uint32_t* verts;
  uint64_t nverts;
  uint64_t nitems;               // row_first[nverts] - row_first[0]
  const uint64_t* row_first;
  const uint64_t* word_first;
  unsigned long long* rows;
  // CliqueRowsKernel with a hash of N+(v) in LDS (d <= kRowsHashMax): vertex
  // t owns chunks [chunk_first[t], chunk_first[t+1]) - chunk_first[0] of 64
  // rows each
  const uint64_t* chunk_first;
  uint64_t nchunks;
  // k == 3, nullable: tc[p] += the triangles through the edge stored at
  // cols[p] (p an index of some N+(x)), recorded by CliqueSmallKernel and
  // CliqueRowsKernel as they find each triangle (v, u_i, u_j)
  uint32_t* tc;
};
// ---------------------------------------------------------------------------
// House (5-cycle with one chord closing a triangle) counted from edge
// statistics (house.hip).  With t(e) the triangles through edge e and p(e)
// the 4-cycles through e, every house is a triangle on its chord e plus a
// 4-cycle through e whose two other vertices avoid the triangle's apex:
//   houses = sum_e t(e) p(e) - 2 sum_e t(e) (t(e) - 1).
// sum_e t(e) p(e) is a sum over 4-cycles of the t of their four edges.  Each
// 4-cycle has one vertex u of highest id (on a degree-ranked graph: highest
// degree) and one vertex w opposite it; with W(u, w) the c middle vertices v
// (v ~ u, v ~ w, v < u), the 4-cycles of (u, w) are the pairs of W and
// together contribute (c - 1) * s, s = sum_{v in W} t(u, v) + t(v, w).  The
// wedges u - v - w (v, w < u) are walked from each apex u and accumulated
// per w as one packed u64 (count << kHouseShift | sum of t).
constexpr int kHouseShift = 40;
constexpr uint32_t kHouseHashMax = 512;  // apex loads a wave hashes in LDS

struct HouseArgs {
  const uint64_t* rp;
  const uint32_t* cols;
  const uint64_t* up;
  uint64_t n;
  uint32_t* t;                  // per CSR entry: triangles through its edge
  unsigned long long* load;     // HouseMirrorKernel: wedges with apex u
  unsigned long long* acc;      // acc[0] += sum (c - 1) s; acc[1] += sum_e t(e) (t(e) - 1)
  // wedge kernels: apex vertices (HouseWedgeHashKernel: load <= kHouseHashMax)
  const uint32_t* apex;
  uint64_t napex;
  // HouseWedgeDenseKernel / HouseWedgeReduceKernel: apex x accumulates in
  // slots[x * n + w]; its items (one per lower neighbour v) are
  // [item_first[x], item_first[x+1]).  The first nfull apexes are reduced by
  // a scan of [0, u); the others through the list of touched w.
  const uint64_t* item_first;
  uint64_t nitems;
  uint64_t nfull;
  unsigned long long* slots;
  uint32_t* touched;            // napex x n
  unsigned long long* ntouched; // napex
};
// Lower CSR entries (v < u) take t from their mirror in N+(v); load[u] and
// acc[1] as above.  The upper entries' t must be complete.
void LaunchHouseMirror(const HouseArgs& a, hipStream_t stream);
// Over the upper entries of every vertex: acc[0] += t, acc[1] += t(t-1)/2
// (the diamonds on that edge: two triangles sharing it).
void LaunchEdgePairs(const HouseArgs& a, hipStream_t stream);
void LaunchHouseWedgeHash(const HouseArgs& a, hipStream_t stream);
void LaunchHouseWedgeDense(const HouseArgs& a, hipStream_t stream);
void LaunchHouseWedgeReduce(const HouseArgs& a, hipStream_t stream);

void LaunchCliqueSmall(const CliqueArgs& a, hipStream_t stream);
void LaunchCliqueRows(const CliqueArgs& a, hipStream_t stream);
// Rows of vertices with d <= kRowsHashMax: a workgroup per chunk of 64 rows
// keeps N+(v) as a hash (value -> local index) in LDS and streams each row
// vertex's upper list, 64 consecutive entries per load; `small_table`: every
// d <= kRowsHashSmall.
constexpr uint32_t kRowsHashMax = 4096;
constexpr uint32_t kRowsHashSmall = 512;
void LaunchCliqueRowsHash(const CliqueArgs& a, bool small_table, hipStream_t stream);
void LaunchCliqueCount(const CliqueArgs& a, hipStream_t stream);

// Exclusive scan: out[i] = sum(in[0..i)), out[n] = total.  `tmp` needs
// ScanTempWords(n) u64 words.  out may alias in.
size_t ScanTempWords(uint64_t n);
void LaunchScanU64(const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* tmp,
                   hipStream_t stream);
void LaunchScanU32(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* tmp,
                   hipStream_t stream);

}  // namespace sme
