// engine.h -- per-device context, resident graph and the level/batch engine.
//
// Context  ~ CudaContext + CudaContextManager (src/gpu/utils/context/
//            CudaContext.cuh:83-98, CudaContextManager.cuh:19-103): one stream
//            and one DevicePool per device.
// DevGraph ~ DevGraphPartition (src/gpu/graph/DevGraphPartition.cuh:32-41).
// Engine   ~ HybridGPUProcessor::ProcessLevel + HybridLIGHTGPUProcessor
//            (src/gpu/impl/common/HybridGPUProcessor.cuh:14-47,
//            src/gpu/impl/light/HybridLIGHTGPUProcessor.cuh:22-316): a
//            depth-first walk over memory-capped batches of each BFS level.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "pool.h"
#include "query.h"

namespace sme {

// Kernel families timed with HIP events when profiling is on.
enum KernelSlot : int {
  kSlotPrep = 0,
  kSlotScan,
  kSlotPartition,
  kSlotIntersect,   // item verify kernels of COMPUTE / CPC / COMPUTE_COUNT
  kSlotLane,        // COMPUTE_COUNT set-up with lane-per-path counting
  kSlotFilter,      // item verify kernels of MATERIALIZE / FILTER_COMPUTE
  kSlotCount,       // COUNT kernels
  kSlotCompact,     // candidate compaction + segment windows
  kSlotExpand,      // MATERIALIZE scatter of partial instances
  kSlotGroup,       // COMPUTE_COUNT of paths sharing a staged segment
  kSlotCliqueSmall, // clique counting, start vertices with <= 64 upper neighbours
  kSlotCliqueRows,  // clique counting, bit rows of larger neighbourhoods
  kSlotCliqueCount, // clique counting over those rows
  kSlotHouseMirror, // house: mirrored triangle counts and wedge loads
  kSlotHouseWedge,  // house: 4-cycle wedges per apex
  kSlotMisc,
  kNumSlots
};
const char* SlotName(int slot);

struct SlotStats {
  uint64_t launches = 0;
  double ms = 0;
  double bytes = 0;  // algorithmic bytes
  uint64_t items = 0;
};

struct RunStats {
  uint64_t batches = 0;
  uint64_t d2h_syncs = 0;
  uint64_t peak_pool_bytes = 0;
  double elapsed_ms = 0;
  uint64_t ranked = 0;  // 1: ran on the degree-ranked copy of the graph
};

class Context {
 public:
  ~Context();
  // Own pool of budget_bytes.
  Status Open(int device, uint64_t budget_bytes, std::string* err);
  // Allocates from `shared` (several contexts on one device, e.g. the slots
  // of a server); `shared` must outlive the context.
  Status OpenShared(int device, DevicePool* shared, std::string* err);
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  DevicePool& pool() { return *pool_; }
  // Runs of this context allocate from `p` from now on; returns the pool
  // used before.
  DevicePool* SwapPool(DevicePool* p) {
    DevicePool* old = pool_;
    pool_ = p;
    return old;
  }

  // Allocation accounting of this context's runs.  With `limit` set (a
  // reservation in a shared pool) the batch planner sees at most
  // limit - held bytes.
  void* Alloc(size_t bytes);
  void Free(void* p, size_t bytes);
  uint64_t RunAvailable() const;
  uint64_t limit = 0;
  uint64_t held = 0;
  uint64_t held_peak = 0;
  uint64_t* pinned() { return pinned_; }

  // hub adjacency bitmaps for graphs uploaded from now on: vertices of
  // degree >= hub_min_degree, highest degree first, at most hub_max_bytes
  // (0 disables)
  uint64_t hub_min_degree = 0;
  uint64_t hub_max_bytes = 0;
  // graphs uploaded from now on also keep a degree-ranked copy (RankByDegree)
  // on which clique plans over all start vertices run
  bool rank_order = false;

  // Schedule tunables (chosen by hand or by the autotuner):
  //   lane_path_max  COMPUTE_COUNT paths with at most this many pivot items
  //                  are counted by one lane in the set-up kernel; longer
  //                  paths go through the 64-items-per-wave schedule (0: all
  //                  paths use the item schedule)
  //   item_grid      workgroup cap of the item kernels (0: default)
  //   super_chunks   chunks per wave in the verify kernels (single-path
  //                  runs take the staged fast path; 0 or 1: off; 4 by
  //                  default, the fastest on R-MAT scale 20)
  //   batch_bytes    cap on the device bytes of one batch (0: the level's
  //                  share of the pool)
  //   stream_ratio   pivot rule for a reused segment plus one adjacency
  //                  list (see PrepComputeArgs; 0: shortest list)
  //   wg_path_min    counting levels: a path with at least this many pivot
  //                  items is counted by a whole workgroup (PathWorkgroupKernel)
  //                  instead of the 64-items-per-wave schedule (0: off)
  //   group_max      counting levels that extend a reused segment with one
  //                  adjacency list: paths sharing a segment of at most this
  //                  many values are counted together by one wave that keeps
  //                  the segment in LDS and streams each path's list
  //                  (GroupCountKernel); 0: off
  //   clique_level_engine  0 (default): clique plans with ordering count
  //                  each start vertex's cliques over bitmaps of its upper
  //                  neighbourhood (clique.hip); 1: through the level engine
  //   census_level_engine  0 (default): a house count over all start vertices
  //                  without level outputs comes from per-edge triangle
  //                  counts and a 4-cycle wedge walk (house.hip); 1: the
  //                  level engine
  struct Tuning {
    uint64_t lane_path_max = 0;
    uint32_t item_grid = 0;
    uint32_t super_chunks = 4;
    uint64_t batch_bytes = 0;
    uint32_t stream_ratio = 0;
    uint32_t group_max = 0;
    uint64_t wg_path_min = 0;
    uint32_t clique_level_engine = 0;
    uint32_t census_level_engine = 0;
  } tuning;

  // profiling
  bool profiling = false;
  std::array<SlotStats, kNumSlots> slots{};
  void ResetProfile() { slots = {}; }
  std::string last_error;

  // event pairs recorded during a run, resolved after the final sync
  struct Mark {
    hipEvent_t start, stop;
    int slot;
    double bytes;
    uint64_t items;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> spare_events;
  hipEvent_t TakeEvent();
  void ResolveMarks();

 private:
  int device_ = -1;
  hipStream_t stream_ = nullptr;
  DevicePool own_pool_;
  DevicePool* pool_ = &own_pool_;
  uint64_t* pinned_ = nullptr;
};

class DevGraph {
 public:
  ~DevGraph();
  // Copies the CSR (plus hub bitmaps and, with ctx->rank_order, the
  // degree-ranked copy as configured on the context).
  Status Upload(Context* ctx, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                uint64_t nnz, std::string* err);
  uint64_t n = 0, nnz = 0;
  uint64_t* rp = nullptr;
  uint32_t* cols = nullptr;
  // up[v] = index in cols of v's first neighbour above v (rp[v + 1] if none)
  uint64_t* up = nullptr;
  // isomorphic degree-ranked copy (RankByDegree), or nullptr
  DevGraph* ranked = nullptr;
  bool self_loops = false;
  uint64_t max_degree = 0;     // largest |N(v)|
  uint64_t max_up_degree = 0;  // largest |N+(v)| (neighbours above v)
  // Built at upload for the clique bitmap path over all start vertices:
  // the vertices with kLocalSmall < |N+(v)| <= kLocalMax in decreasing
  // |N+(v)| (ties by id) with prefix sums of their rows, row words and
  // 64-row chunks (host and device copies), and those above kLocalMax.
  struct UpperLists {
    std::vector<uint32_t> verts, d, huge;
    std::vector<uint64_t> rf, wf, cf;
    uint32_t* dverts = nullptr;
    uint64_t* drf = nullptr;
    uint64_t* dwf = nullptr;
    uint64_t* dcf = nullptr;
  } upper;
  int device = -1;
  // hub adjacency bitmaps (nhub rows of hub_words u32 words)
  uint32_t* hub_id = nullptr;
  uint32_t* hub_bits = nullptr;
  uint64_t nhub = 0, hub_words = 0;
};

// Runs one plan over start vertices [v_begin, v_end) of the first matched
// vertex (an IntraPartTask, src/gpu/task/Task.h:23-43).  A run over all
// start vertices goes to the degree-ranked copy when the graph has one and
// the answer cannot differ there: a clique plan, or no level outputs asked
// for.  With `ranked_ids` the range is in the ranked copy's ids (a task of a
// partition of [0, n) over that copy; the copy must exist).
Status RunPlan(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t v_begin,
               uint32_t v_end, uint64_t* count, uint64_t* level_counts, RunStats* stats,
               std::string* err, bool ranked_ids = false);

// Same walk over [bounds[0], bounds[nbuckets]); additionally counts[k]
// receives the embeddings whose first matched vertex lies in
// [bounds[k], bounds[k+1]).  bounds must be non-decreasing.
Status RunPlanBuckets(Context* ctx, const DevGraph* g, const EnginePlan* plan,
                      const uint32_t* bounds, uint64_t nbuckets, uint64_t* counts, uint64_t* count,
                      uint64_t* level_counts, RunStats* stats, std::string* err);

// The no-allocation variant of RunPlan: every intermediate comes from the
// device's static scratch array (static_scratch.hip) instead of the
// context's pool, and no HIP events are created.  Calls that overlap on
// one device share that array and race.
Status RunPlanNoAlloc(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t v_begin,
                      uint32_t v_end, uint64_t* count, uint64_t* level_counts, RunStats* stats,
                      std::string* err);
// Once per device, with the device current: locates the static scratch
// array (hipGetSymbolAddress) and sets up its pool.  Called by Context::Open.
Status InitStaticScratch(int device, std::string* err);
uint64_t StaticScratchBytes();

// Match list of a full-materialization plan (GPSM / GPSM_REUSE) over start
// vertices [v_begin, v_end): up to max_rows tuples of plan->nv u32 written
// to host memory `out` in lexicographic order; *total = the full count,
// *rows = tuples written.
Status ListMatches(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t v_begin,
                   uint32_t v_end, uint32_t* out, uint64_t max_rows, uint64_t* total, uint64_t* rows,
                   std::string* err);

}  // namespace sme
