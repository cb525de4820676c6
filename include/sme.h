/*
 * sme.h -- C ABI of the MI355X subgraph enumeration engine (libsme.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an sme_status
 * (0 = OK) and never exits the process; the message of the last failure of a
 * context is available from sme_last_error().  A context is used from one
 * host thread at a time (one context per device, like the reference's one
 * host thread per device, src/gpu/executor/HybridSeparatePatternMatch.cuh:78).
 *
 * Reference interfaces each entry point replaces (paths in the reference
 * tree, src/...):
 *   sme_open / sme_close        CudaContextManager::CreateCudaContextManager +
 *                               CnmemManagedCudaContext (gpu/utils/context/
 *                               CudaContextManager.cuh:19-103,
 *                               CudaContext.cuh:315-338) and the device
 *                               memory limit kDeviceMemoryLimits (Meta.h:26-28)
 *   sme_graph_upload            HybridGPUComponent::BuildDevicePartition ->
 *                               DevGraphPartition::BuildIntraPartition
 *                               (gpu/impl/common/HybridGPUComp.cuh:87-94,
 *                               gpu/graph/DevGraphPartition.cuh:32-41)
 *   sme_graph_destroy           HybridGPUComponent::ReleaseDevicePartition
 *                               (HybridGPUComp.cuh:83-86)
 *   sme_plan_create             InitGPUInstance plan selection
 *                               (gpu/PatternMatchGPUInstance.cuh:120-146):
 *                               Query + LazyTraversalPlan / TraversalPlan
 *                               (query/common/Query.h:13-53,
 *                               query/LazyTraversalPlan.h:13-29)
 *   sme_plan_describe           LazyTraversalPlan::Print (LazyTraversalPlan.h:68-86)
 *   sme_count                   HybridGPUComponent::GPUThreadExecute(IntraPartTask)
 *                               (HybridGPUComp.cuh:53-69, task/Task.h:23-43)
 *   sme_bin_read / sme_bin_write  GraphIO::ReadCSRBinFile / WriteCSRBinFile
 *                               (graph/GraphIO.h:88-136, 356-387)
 *   sme_snap_read               GraphIO::ReadSNAPFile (GraphIO.h:232-311)
 *   sme_gen_rmat / sme_gen_er   no counterpart (the reference reads SNAP
 *                               datasets); seeded synthetic inputs
 *   sme_profile_*               GPUProfiler phase timers
 *                               (gpu/utils/GPUProfiler.cuh:11-134)
 */
#ifndef SME_H_
#define SME_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sme_ctx sme_ctx;
typedef struct sme_graph sme_graph;
typedef struct sme_plan sme_plan;

enum sme_status {
  SME_OK = 0,
  SME_ERR_INVALID = 1,
  SME_ERR_OOM = 2,
  SME_ERR_HIP = 3,
  SME_ERR_IO = 4,
  SME_ERR_UNSUPPORTED = 5
};

/* GPU algorithms, numbered as Algo in the reference (Meta.h:47-66). */
enum sme_algo { SME_ALGO_GPSM = 4, SME_ALGO_GPSM_REUSE = 5, SME_ALGO_LIGHT = 6 };

/* LazyTraversalCompressLevel (query/common/LazyTraversalCommon.h:38-76). */
enum sme_compress_level {
  SME_CL_MATERIALIZE = 0,
  SME_CL_NON_MATERIALIZE = 1,
  SME_CL_NON_MATERIALIZE_OPT = 2,
  SME_CL_SPECIAL = 3
};

#define SME_MAX_LEVELS 16

typedef struct {
  uint64_t batches;          /* level batches executed */
  uint64_t d2h_syncs;        /* blocking device->host reads */
  uint64_t peak_pool_bytes;  /* high-water mark of the intermediate pool */
  double elapsed_ms;         /* host wall time of the call */
  uint64_t ranked;           /* 1: the call ran on the degree-ranked copy */
} sme_stats;

/* ---- context ---------------------------------------------------------- */
/* Opens `device` and reserves `mem_budget_bytes` of device memory for
 * intermediates (the reference's per-device limit, Meta.h:26-28). */
int sme_open(int device, uint64_t mem_budget_bytes, sme_ctx** out);
void sme_close(sme_ctx* ctx);
const char* sme_last_error(const sme_ctx* ctx);
int sme_device_count(int* out);

/* ---- graph ------------------------------------------------------------ */
/* Copies a CSR (BIN layout: row_ptrs[n+1] u64, cols[nnz] u32, every list
 * strictly increasing) to the context's device.  The caller keeps its
 * arrays. */
int sme_graph_upload(sme_ctx* ctx, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                     uint64_t nnz, sme_graph** out);
void sme_graph_destroy(sme_graph* g);
/* Adjacency bitmaps for high-degree vertices, built by later uploads: every
 * vertex of degree >= min_degree, highest degree first, within max_bytes of
 * device memory (one bit per vertex id per row).  Membership tests against
 * those lists become one bit test instead of a binary search.  0 disables
 * (the default).  No counterpart in the reference, which always
 * binary-searches (gpu/utils/ops/GPUBinSearch.cuh:18-31). */
int sme_set_hub_bitmaps(sme_ctx* ctx, uint64_t min_degree, uint64_t max_bytes);
int sme_graph_hubs(const sme_graph* g, uint64_t* nhub);
/* Degree-ranked copy, built by later uploads when on: the same graph with
 * vertex ids reassigned in increasing (degree, id) order and every list
 * re-sorted (one more CSR on the device).  sme_count over all start
 * vertices runs on the copy when the answer cannot differ: a clique plan
 * (Q0, Q3, Q7, Q11), whose count and per-level outputs are all invariant
 * under relabeling, or any plan when level_counts is NULL (a count is
 * invariant).  Listing each instance from its lowest-degree vertex bounds
 * the lists walked (Chiba-Nishizeki orientation).  Start-range, bucketed
 * and match-list calls use the original ids.  Off by default.  No
 * counterpart in the reference, which lists from the lowest id
 * (LazyTraversalPlan.h:27-28). */
int sme_set_rank_order(sme_ctx* ctx, int on);
int sme_graph_ranked(const sme_graph* g, int* on);

/* ---- plan (host only) ------------------------------------------------- */
int sme_plan_create(int algo, int query_type, int compress_level, int enable_ordering,
                    sme_plan** out);
/* Canonical text of the plan; returns the full length (may exceed cap). */
int sme_plan_describe(const sme_plan* p, char* buf, size_t cap);
/* Number of exec levels (length of the level_counts array of sme_count). */
int sme_plan_levels(const sme_plan* p);
/* Intersection reuse across levels (on by default for LIGHT and
 * GPSM_REUSE, off for GPSM): a level whose backward lists include the
 * previous vertex's lists extends that vertex's candidate segment instead of
 * re-intersecting adjacency lists, where the conditions prove the result is
 * the same set.  Counterpart of the reuse plans, query/ReuseTraversalPlan.h:9-77. */
int sme_plan_set_reuse(sme_plan* p, int on);
void sme_plan_destroy(sme_plan* p);

/* ---- execution -------------------------------------------------------- */
/* Counts the embeddings whose first matched vertex lies in [v_begin, v_end).
 * level_counts (nullable) receives, per exec level, the summed output size:
 * partial instances after MATERIALIZE, candidates after COMPUTE /
 * FILTER_COMPUTE / COMPUTE_PATH_COUNT, the count added by COUNT /
 * COMPUTE_COUNT. */
int sme_count(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t v_begin,
              uint32_t v_end, uint64_t* out_count, uint64_t* level_counts, sme_stats* stats);

/* sme_count that allocates nothing during the call.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory while
 * they count (code under stream capture, tight serving loops).  Every
 * intermediate comes from a device array of sme_static_scratch_bytes()
 * (64 MiB) that the library holds on each device it runs on until the
 * process ends (set up at sme_open), instead of the context's pool; no
 * hipMalloc, hipFree or HIP event is created by the call, and the batches
 * shrink to fit the array.  A call whose single unsplittable batch needs
 * more than the array is rejected with SME_ERR_OOM and a message naming the
 * limit.  All calls on one device share the array: calls that overlap on
 * one device race.  Results equal sme_count's. */
int sme_count_static(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t v_begin,
                     uint32_t v_end, uint64_t* out_count, uint64_t* level_counts,
                     sme_stats* stats);
uint64_t sme_static_scratch_bytes(void);

/* sme_count over a start range [r_begin, r_end) of the graph's degree-ranked
 * copy (sme_set_rank_order), in that copy's ids: the tasks of any partition
 * of [0, n) sum to the full count (and, for cliques, to the full level
 * outputs), so a long count can run as several calls.  SME_ERR_INVALID when
 * the graph has no ranked copy.  No counterpart in the reference; it plays
 * the role of IntraPartTask ranges (src/gpu/task/Task.h:23-43) on the
 * relabeled graph. */
int sme_count_ranked(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t r_begin,
                     uint32_t r_end, uint64_t* out_count, uint64_t* level_counts,
                     sme_stats* stats);

/* One enumeration over [bounds[0], bounds[nbuckets]) that also splits the
 * count by start vertex: counts[k] receives the embeddings whose first
 * matched vertex lies in [bounds[k], bounds[k+1]) (bounds non-decreasing).
 * Used for graphs packed back to back (one bucket per graph) and for merged
 * requests over one resident graph.  No counterpart in the reference, which
 * sums one count per task (src/gpu/executor/HybridSeparatePatternMatch.cuh:91-95). */
int sme_count_buckets(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, const uint32_t* bounds,
                      uint64_t nbuckets, uint64_t* counts, uint64_t* out_total,
                      uint64_t* level_counts, sme_stats* stats);

/* Match list of a full-materialization plan (SME_ALGO_GPSM or
 * SME_ALGO_GPSM_REUSE) over start vertices [v_begin, v_end): each match is a
 * row of sme_plan_vertices(p) u32, the data vertex of query vertex 0, 1, ...,
 * and rows come in lexicographic order (the order GPSM's levels produce
 * them).  At most max_matches rows are written to out; *total receives the
 * full count and *written the rows written.  A LIGHT plan is
 * SME_ERR_UNSUPPORTED (it counts without materializing every vertex).  No
 * counterpart in the reference: its -l flag is parsed
 * (src/gpu/PatternMatchGPU.cu:29) but never consumed. */
int sme_list_matches(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t v_begin,
                     uint32_t v_end, uint32_t* out, uint64_t max_matches, uint64_t* total,
                     uint64_t* written);
/* Query vertices of the plan's pattern (row width of sme_list_matches). */
int sme_plan_vertices(const sme_plan* p);

/* ---- schedule tuning ---------------------------------------------------
 * No counterpart in the reference, whose launch shapes are compile-time
 * constants (gpu/utils/ops/GPUIntersection.cuh, LIGHTCommon.cuh).
 *   lane_path_max  counting levels (COMPUTE_COUNT): a path whose pivot list
 *                  window holds at most this many candidates is verified by
 *                  one lane of the set-up kernel; longer paths go through the
 *                  64-candidates-per-wavefront schedule.  0: every path uses
 *                  the wavefront schedule (the default).
 *   item_grid      workgroup cap of the candidate kernels (0: 16384); every
 *                  wavefront loops over 64-candidate chunks, so a cap near
 *                  the resident capacity makes the launch persistent.
 *   super_chunks   64-candidate chunks a wavefront takes at once; when they
 *                  all belong to one path, that path's lists are loaded once
 *                  and one short list is staged in LDS (0 or 1: chunk by
 *                  chunk; a new context starts at 4).
 *   batch_bytes    cap on the device bytes one batch of a level may use
 *                  (0: the level's share of the memory budget).
 *   stream_ratio   pivot rule when a level extends a reused candidate
 *                  segment with one adjacency list: if the segment holds at
 *                  most 1024 candidates and the list at most stream_ratio
 *                  times as many, the list is read in order as the pivot and
 *                  the segment is probed (from LDS) instead of probing the
 *                  list at random (0: the shorter is the pivot).
 *   group_max      counting levels that extend a reused candidate segment
 *                  with one adjacency list: the paths sharing a segment of
 *                  at most group_max values (capped at 512) are counted by
 *                  one wavefront that keeps the segment in LDS and reads
 *                  each path's list in order, 64 consecutive entries at a
 *                  time (0: off, the default).
 *   wg_path_min    counting levels: a path with at least this many pivot
 *                  candidates is counted by a whole workgroup (256 threads
 *                  that stage a window of another list in LDS, up to 8,192
 *                  entries) instead of the 64-candidates-per-wavefront
 *                  schedule; the reference gives every path one warp
 *                  (GPSMBinSearch.cuh:36-63).  0: off (the default).
 *   clique_level_engine  0 (the default): a LIGHT clique plan with ordering
 *                  (triangle, 4-, 5-, 6-clique) counts the cliques of each
 *                  start vertex v inside its upper neighbourhood N+(v) by
 *                  AND + popcount over bit rows of that neighbourhood; 1:
 *                  through the level engine like every other plan.  Counts
 *                  and level outputs are identical.
 *   census_level_engine  0 (the default): a house count (Q4, ordering on)
 *                  over all start vertices with no level outputs asked for
 *                  is computed from per-edge triangle counts and a
 *                  degree-ordered 4-cycle walk: houses = sum_e t(e) p(e) -
 *                  2 sum_e t(e)(t(e)-1), t(e) and p(e) the triangles and
 *                  4-cycles through edge e (graphs without self-loops, every
 *                  degree below 2^19); 1: through the level engine.  The
 *                  count is identical.
 * Settings apply to every later call on the context; results never depend
 * on them. */
typedef struct {
  uint64_t lane_path_max;
  uint32_t item_grid;
  uint32_t super_chunks;
  uint64_t batch_bytes;
  uint32_t stream_ratio;
  uint32_t group_max;
  uint64_t wg_path_min;
  uint32_t clique_level_engine;
  uint32_t census_level_engine;
} sme_tuning;
int sme_set_tuning(sme_ctx* ctx, const sme_tuning* t);
int sme_get_tuning(const sme_ctx* ctx, sme_tuning* t);

typedef struct {
  sme_tuning tuning;
  double ms; /* summed time of the sample runs */
} sme_tune_trial;

/* Autotuner: times the plans on a sample of the start vertices for each
 * candidate setting (coordinate descent over lane_path_max, super_chunks,
 * item_grid, batch_bytes, stream_ratio, group_max and wg_path_min, starting
 * from the context's current setting; every setting's sample runs three
 * times and its median counts; a setting replaces the current one only when
 * at least 2 % faster), installs the fastest on the context and returns it.  With bounds (nbuckets >= 1) the
 * sample is 8 evenly spaced blocks of consecutive buckets covering about
 * sample_fraction of the buckets, run through sme_count_buckets; without
 * (bounds == NULL) it is 8 evenly spaced start-vertex ranges covering
 * sample_fraction of [0, n).  trials (nullable, cap entries) receives every
 * measured setting, *ntrials their number. */
int sme_autotune(sme_ctx* ctx, const sme_graph* g, const sme_plan* const* plans, int nplans,
                 const uint32_t* bounds, uint64_t nbuckets, double sample_fraction,
                 sme_tuning* best, double* best_ms, sme_tune_trial* trials, int cap,
                 int* ntrials);

/* ---- server: pattern-count requests over one resident graph ------------
 * No counterpart in the reference (one pattern per process run,
 * src/gpu/PatternMatchGPU.cu:17-37).  Requests are (query, start-vertex
 * range) pairs, counted as sme_count would count them (LIGHT, SPECIAL
 * compress level, ordering on).  `slots` passes run at once, each on its own
 * stream, sharing one device pool of mem_cap_bytes.
 *   merge != 0: continuous batching -- when a slot frees up, every queued
 *     request of the oldest request's pattern (at most max_merge) becomes
 *     one pass over the union of their ranges, split per request by
 *     per-start-vertex buckets.
 *   admission: a pass starts when its estimated intermediate bytes fit
 *     beside the passes running (estimate = learned bytes per work unit of
 *     the pattern x summed squared degrees of the covered start vertices,
 *     clamped to [min_reserve_bytes, mem_cap_bytes]; a pattern's first pass
 *     reserves mem_cap_bytes / slots).
 * merge = 0 with one slot is the plain path: one request at a time, in
 * arrival order, with the whole budget.  Submit and wait may be called from
 * any host thread. */
typedef struct sme_server sme_server;
typedef struct {
  uint64_t mem_cap_bytes;
  int32_t slots;
  int32_t merge;
  uint64_t max_merge;
  uint64_t min_reserve_bytes;
  uint64_t hub_min_degree; /* adjacency bitmaps, as sme_set_hub_bitmaps */
  uint64_t hub_max_bytes;
  sme_tuning tuning;       /* schedule of every pass */
} sme_server_config;
typedef struct {
  uint64_t passes;          /* engine passes run */
  uint64_t requests_done;
  uint64_t max_group;       /* most requests one pass served */
  uint64_t peak_reserved;   /* largest sum of admitted reservations (bytes) */
  uint64_t peak_concurrent; /* most passes running at once */
  uint64_t failed;
} sme_server_stats;
int sme_server_open(int device, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                    uint64_t nnz, const sme_server_config* cfg, sme_server** out);
int sme_server_submit(sme_server* s, int query, uint32_t v_begin, uint32_t v_end, uint64_t* id);
/* Blocks until request `id` is done; latency_ms = submit to completion. */
int sme_server_wait(sme_server* s, uint64_t id, uint64_t* count, double* latency_ms);
int sme_server_get_stats(sme_server* s, sme_server_stats* out);
/* Finishes every submitted request, then releases the server. */
void sme_server_close(sme_server* s);

/* ---- kernel profile (HIP events around every launch) ------------------ */
int sme_profile_enable(sme_ctx* ctx, int on);
int sme_profile_reset(sme_ctx* ctx);
int sme_profile_slots(void);
/* Per kernel family: name, launches, summed ms, algorithmic bytes, items. */
int sme_profile_get(const sme_ctx* ctx, int slot, char* name, size_t cap, uint64_t* launches,
                    double* ms, double* bytes, uint64_t* items);

/* Diagnostic with no counterpart in the reference: copies `bytes` within the
 * context's pool with one device-to-device copy and returns its time.  Used
 * to calibrate hardware traffic counters against a copy of known size. */
int sme_calibrate_copy(sme_ctx* ctx, uint64_t bytes, double* ms);

/* ---- host graph utilities (no device needed) --------------------------- */
/* Arrays returned through row_ptrs / cols are released with sme_free_host. */
int sme_bin_read(const char* path, uint64_t* n, uint64_t* nnz, uint64_t** row_ptrs,
                 uint32_t** cols);
int sme_bin_write(const char* path, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                  uint64_t nnz);
int sme_snap_read(const char* path, int directed, uint64_t* n, uint64_t* nnz, uint64_t** row_ptrs,
                  uint32_t** cols);
int sme_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 uint64_t* n, uint64_t* nnz, uint64_t** row_ptrs, uint32_t** cols);
int sme_gen_er(uint64_t n_vertices, uint64_t n_edges, uint64_t seed, uint64_t* n, uint64_t* nnz,
               uint64_t** row_ptrs, uint32_t** cols);
/* `ngraphs` seeded G(n, m) graphs packed back to back into one CSR (a
 * disjoint union): member k has n_k uniform in [n_min, n_max] vertices and
 * average degree uniform in [deg_min, deg_max); its vertices are
 * [graph_offs[k], graph_offs[k+1]) (ngraphs + 1 offsets, released with
 * sme_free_host).  Counting with sme_count_buckets over graph_offs gives
 * every member's count in one pass. */
int sme_gen_er_batch(uint64_t ngraphs, uint64_t n_min, uint64_t n_max, double deg_min,
                     double deg_max, uint64_t seed, uint64_t* n, uint64_t* nnz,
                     uint64_t** row_ptrs, uint32_t** cols, uint64_t** graph_offs);
void sme_free_host(void* p);

#ifdef __cplusplus
}
#endif
#endif /* SME_H_ */
