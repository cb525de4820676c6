// query.h -- pattern registry and the LIGHT / GPSM execution plans.
//
// Mirrors the reference's host-side query layer:
//   QueryType ids Q0..Q13             src/Meta.h:104-124
//   Query (adjacency + symmetry-breaking orders)
//                                     src/query/common/Query.h:11-53
//   LazyTraversalPlan (LIGHT plan)    src/query/LazyTraversalPlan.h:11-198
//   TraversalPlan orderings for GPSM  src/query/TraversalPlan.h:95-103,
//                                     src/query/common/Plan.h:60-71,163-174
// The plan is produced once on the host and handed to the engine as a small
// POD (EnginePlan) that is passed to kernels by value.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "common.h"

namespace sme {

// CondOperator (src/Meta.h:126).
enum CondOp : uint8_t { kLess = 0, kLarger = 1, kNonEqual = 2, kNoOp = 3 };

// LazyTraversalOperation (src/query/common/LazyTraversalCommon.h:11-31).
enum OpKind : uint8_t {
  kCompute = 0,
  kComputePathCount = 1,
  kMaterialize = 2,
  kFilterCompute = 3,
  kCount = 4,
  kComputeCount = 5,
};

// LazyTraversalCompressLevel (LazyTraversalCommon.h:38-76).
enum CompressLevel : int {
  kClMaterialize = 0,
  kClNonMaterialize = 1,
  kClNonMaterializeOpt = 2,
  kClSpecial = 3,
};

// Algo ids of the GPU methods (src/Meta.h:47-66).
enum AlgoId : int { kAlgoGpsm = 4, kAlgoGpsmReuse = 5, kAlgoLight = 6 };

struct Cond {
  CondOp op;
  uint8_t v;
};
using CondList = std::vector<Cond>;
using VList = std::vector<int>;

class Query {
 public:
  // Returns false for an unknown id.  With enable_ordering=false all
  // symmetry-breaking orders are dropped (Query.h:77-82).
  bool Init(int query_type, bool enable_ordering);
  int Type() const { return type_; }
  bool OrderingEnabled() const { return enable_ordering_; }
  int VertexCount() const { return nv_; }
  const std::vector<VList>& Connectivity() const { return con_; }
  const std::vector<CondList>& Ordering() const { return order_; }
  // Plan::GetConditionType (Plan.h:32-43): NON_EQUAL unless ordered.
  CondOp ConditionType(int a, int b) const;

 private:
  int type_ = -1;
  bool enable_ordering_ = true;
  int nv_ = 0;
  std::vector<VList> con_;
  std::vector<CondList> order_;
};

struct OpEntry {
  OpKind op;
  int u;  // query vertex (== vertex count for COUNT)
};

// Host-side LIGHT plan with the same accessors the reference's executor uses.
class LightPlan {
 public:
  bool Build(const Query& q, int compress_level);
  const Query& GetQuery() const { return q_; }
  int CompressLevel() const { return cl_; }
  const std::vector<int>& SearchSequence() const { return seq_; }
  const std::vector<OpEntry>& ExecuteOperations() const { return ops_; }
  const std::vector<VList>& MaterializedVertices() const { return mat_; }
  const std::vector<VList>& ComputedUnmaterializedVertices() const { return cun_; }
  const std::vector<VList>& BackwardConnectivity() const { return bconn_; }
  const std::vector<CondList>& ComputeCondition() const { return ccond_; }
  const std::vector<CondList>& MaterializeCondition() const { return mcond_; }
  const std::vector<CondList>& FilterCondition() const { return fcond_; }
  std::string Describe() const;

 private:
  void ComputeSequence();
  void InspectUnmaterialized(VList* um, std::vector<bool>* need_filter) const;
  void Count1(const VList& um, const std::vector<bool>& nf, int level);
  void Count2(const VList& um, const std::vector<bool>& nf, int level);
  void Constraints();
  size_t OpIndex(OpKind op, int u) const;

  Query q_;
  int cl_ = kClSpecial;
  std::vector<int> seq_;
  std::vector<OpEntry> ops_;
  std::vector<VList> mat_, cun_, bconn_;
  std::vector<CondList> ccond_, mcond_, fcond_;
};

// ---------------------------------------------------------------------------
// Engine-facing POD plan.  One OpDesc per exec level.
struct OpDesc {
  uint8_t kind;       // OpKind
  uint8_t u;          // target query vertex
  uint8_t nb;         // backward lists intersected (compute kinds)
  uint8_t bconn[kMaxQueryVertices];
  uint8_t nc;         // conditions of this op (compute / materialize / filter)
  Cond cond[kMaxQueryVertices];
  uint8_t nmat;       // vertices materialized before this level
  uint8_t mat[kMaxQueryVertices];
  uint8_t ncun;       // computed-but-unmaterialized before this level
  uint8_t cun[kMaxQueryVertices];
  // COMPUTE_COUNT with one other unmaterialized vertex / COUNT with two:
  // operator between the pair (Plan::GetConditionType).
  uint8_t pair_op;
};

struct EnginePlan {
  int algo;            // AlgoId
  int query_type;
  int compress_level;
  int enable_ordering;
  int nv;
  int nops;
  int direct_count;    // COUNT uses the per-path closed form (LIGHTDirectCount)
  int reuse;           // extend the parent's candidate set where provably equal
  // The pattern is a clique: every level's output (partial instances and
  // candidates alike) counts the injective maps of a clique prefix that
  // satisfy conditions symmetric under any vertex relabeling, so a run over
  // all start vertices gives the same count and level outputs on any
  // isomorphic copy of the graph.
  int clique;
  OpDesc ops[kMaxOps];
};

// LIGHT plan -> engine plan.
bool MakeLightEnginePlan(int query_type, int compress_level, bool enable_ordering,
                         EnginePlan* out, std::string* describe);
// GPSM: identity order, GetOrderedConnectivity + GetWholeOrderedOrdering,
// every level computed then materialized; count = size of the last level
// (src/gpu/impl/gpsm/HybridGPSMGPUProcessor.cuh:23-33).  With `reuse` the
// plan is GPSM_REUSE: a level extends the previous level's intersection
// where that is provably the same set (src/gpu/impl/gpsm/reuse).
bool MakeGpsmEnginePlan(int query_type, bool enable_ordering, bool reuse, EnginePlan* out,
                        std::string* describe);

// CanDirectCount (src/gpu/impl/light/op/LIGHTDirectCount.cuh:13-19).
bool CanDirectCount(int query_type);

}  // namespace sme
