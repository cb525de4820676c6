// query.cpp -- pattern registry and plan construction (see query.h).
#include "query.h"

#include <algorithm>
#include <cstdio>

namespace sme {

namespace {

// One registry row: per-vertex adjacency in the reference's insertion order
// and per-vertex symmetry-breaking orders (operator, operand).  The insertion
// order fixes the order of backward neighbours, which the LIGHT plan
// generator uses to decide which vertices to materialize first.
struct PatternRow {
  int nv;
  std::vector<VList> adj;
  std::vector<CondList> ord;
};

Cond Lt(int v) { return Cond{kLess, static_cast<uint8_t>(v)}; }
Cond Gt(int v) { return Cond{kLarger, static_cast<uint8_t>(v)}; }

// src/query/common/Query.h:84-433
const std::vector<PatternRow>& Registry() {
  static const std::vector<PatternRow> rows = {
      // Q0 triangle
      {3, {{1, 2}, {0, 2}, {0, 1}}, {{Lt(1)}, {Gt(0), Lt(2)}, {Gt(1)}}},
      // Q1 square (4-cycle)
      {4, {{1, 3}, {0, 2}, {1, 3}, {0, 2}},
       {{Lt(1), Lt(2), Lt(3)}, {Gt(0), Lt(3)}, {Gt(0)}, {Gt(0), Gt(1)}}},
      // Q2 chordal square (diamond)
      {4, {{1, 2, 3}, {0, 2}, {0, 1, 3}, {0, 2}}, {{Lt(2)}, {Lt(3)}, {Gt(0)}, {Gt(1)}}},
      // Q3 4-clique
      {4, {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}},
       {{Lt(1)}, {Gt(0), Lt(2)}, {Gt(1), Lt(3)}, {Gt(2)}}},
      // Q4 house
      {5, {{1, 2}, {0, 2, 3}, {0, 1, 4}, {1, 4}, {2, 3}}, {{}, {Lt(2)}, {Gt(1)}, {}, {}}},
      // Q5 quad triangle
      {6, {{1, 2, 3, 4, 5}, {0, 2}, {0, 1, 3}, {0, 2, 4}, {0, 3, 5}, {0, 4}},
       {{}, {}, {Lt(4)}, {}, {Gt(2)}, {}}},
      // Q6 near-5-clique
      {5, {{1, 2, 3, 4}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2, 4}, {0, 3}},
       {{Lt(3)}, {Lt(2)}, {Gt(1)}, {Gt(0)}, {}}},
      // Q7 5-clique
      {5, {{1, 2, 3, 4}, {0, 2, 3, 4}, {0, 1, 3, 4}, {0, 1, 2, 4}, {0, 1, 2, 3}},
       {{Lt(1)}, {Gt(0), Lt(2)}, {Gt(1), Lt(3)}, {Gt(2), Lt(4)}, {Gt(3)}}},
      // Q8 chordal roof
      {5, {{1, 2, 3}, {0, 2, 3}, {0, 1, 4}, {0, 1, 4}, {2, 3}},
       {{Lt(1)}, {Gt(0)}, {Lt(3)}, {Gt(2)}, {}}},
      // Q9 three triangle
      {5, {{1, 2, 3, 4}, {0, 3, 2}, {0, 1, 4}, {0, 1}, {0, 2}}, {{}, {Lt(2)}, {Gt(1)}, {}, {}}},
      // Q10 solar square
      {5, {{1, 2, 3, 4}, {0, 2, 4}, {0, 1, 3}, {0, 2, 4}, {0, 1, 3}},
       {{}, {Lt(2), Lt(3), Lt(4)}, {Lt(4), Gt(1)}, {Gt(1)}, {Gt(1), Gt(2)}}},
      // Q11 6-clique
      {6,
       {{1, 2, 3, 4, 5}, {0, 2, 3, 4, 5}, {0, 1, 3, 4, 5}, {0, 1, 2, 4, 5}, {0, 1, 2, 3, 5},
        {0, 1, 2, 3, 4}},
       {{Lt(1)}, {Gt(0), Lt(2)}, {Gt(1), Lt(3)}, {Gt(2), Lt(4)}, {Gt(3), Lt(5)}, {Gt(4)}}},
      // Q12 near-near-5-clique
      {5, {{1, 2, 3, 4}, {0, 2, 3, 4}, {0, 1, 3}, {0, 1, 2, 4}, {0, 1, 3}},
       {{Lt(1)}, {Gt(0)}, {Lt(4)}, {}, {Gt(2)}}},
      // Q13 triangle in diamond
      {6, {{1, 2, 3, 5}, {0, 2, 3, 4}, {0, 1, 4, 5}, {0, 1}, {1, 2}, {0, 2}},
       {{Lt(1)}, {Lt(2), Gt(0)}, {Gt(1)}, {}, {}, {}}},
  };
  return rows;
}

// Hard-coded LIGHT match orders (LazyTraversalPlanHardCode.h:59-117).
const std::vector<VList>& LightOrders() {
  static const std::vector<VList> orders = {
      {0, 1, 2},       {0, 1, 2, 3},    {0, 2, 1, 3},       {0, 1, 2, 3},
      {1, 2, 3, 0, 4}, {0, 2, 4, 1, 3, 5}, {0, 1, 3, 2, 4}, {0, 1, 2, 3, 4},
      {0, 1, 2, 3, 4}, {0, 1, 2, 3, 4}, {0, 1, 3, 2, 4},    {0, 1, 2, 3, 4, 5},
      {0, 1, 3, 2, 4}, {0, 1, 2, 3, 4, 5}};
  return orders;
}

const char* OpShort(OpKind k) {
  switch (k) {
    case kCompute: return "C";
    case kComputePathCount: return "CPC";
    case kMaterialize: return "M";
    case kFilterCompute: return "FC";
    case kCount: return "COUNT";
    case kComputeCount: return "CC";
  }
  return "?";
}

void AppendV(std::string* s, const VList& v) {
  *s += "[";
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) *s += ",";
    *s += std::to_string(v[i]);
  }
  *s += "]";
}

void AppendC(std::string* s, const CondList& c) {
  static const char kSym[] = {'<', '>', '!', '?'};
  *s += "[";
  for (size_t i = 0; i < c.size(); ++i) {
    if (i) *s += ",";
    *s += kSym[c[i].op];
    *s += std::to_string(c[i].v);
  }
  *s += "]";
}

}  // namespace

bool Query::Init(int query_type, bool enable_ordering) {
  const auto& reg = Registry();
  if (query_type < 0 || query_type >= static_cast<int>(reg.size())) return false;
  const PatternRow& r = reg[query_type];
  type_ = query_type;
  enable_ordering_ = enable_ordering;
  nv_ = r.nv;
  con_ = r.adj;
  order_.assign(nv_, CondList());
  if (enable_ordering) order_ = r.ord;
  return true;
}

CondOp Query::ConditionType(int a, int b) const {
  for (const Cond& c : order_[a])
    if (c.v == b) return c.op;
  return kNonEqual;
}

size_t LightPlan::OpIndex(OpKind op, int u) const {
  for (size_t i = 0; i < ops_.size(); ++i)
    if (ops_[i].op == op && ops_[i].u == u) return i;
  return ops_.size();
}

// Backward connectivity by match order (Plan.h:202-217) and the COMPUTE /
// MATERIALIZE skeleton (LazyTraversalPlanGenerator.h:124-154): each vertex in
// match order is computed right after its not-yet-materialized backward
// neighbours are materialized.
void LightPlan::ComputeSequence() {
  const int nv = q_.VertexCount();
  bconn_.assign(nv, VList());
  std::vector<bool> seen(nv, false);
  for (int u : seq_) {
    for (int w : q_.Connectivity()[u])
      if (seen[w]) bconn_[u].push_back(w);
    seen[u] = true;
  }
  std::vector<bool> done(nv, false);
  ops_.push_back({kMaterialize, seq_[0]});
  done[seq_[0]] = true;
  for (size_t i = 1; i < seq_.size(); ++i) {
    const int u = seq_[i];
    for (int w : bconn_[u]) {
      if (!done[w]) {
        ops_.push_back({kMaterialize, w});
        done[w] = true;
      }
    }
    ops_.push_back({kCompute, u});
  }
}

// LazyTraversalPlanGenerator.h:158-184: vertices never materialized, and
// whether anything is materialized after their COMPUTE.
void LightPlan::InspectUnmaterialized(VList* um, std::vector<bool>* need_filter) const {
  std::vector<bool> done(q_.VertexCount(), false);
  for (const OpEntry& e : ops_)
    if (e.op == kMaterialize) done[e.u] = true;
  for (int u = 0; u < q_.VertexCount(); ++u) {
    if (done[u]) continue;
    size_t at = OpIndex(kCompute, u);
    bool later = false;
    for (size_t i = at + 1; i < ops_.size(); ++i) later = later || ops_[i].op == kMaterialize;
    um->push_back(u);
    need_filter->push_back(later);
  }
}

// GenerateCount1ExecuteSequence (LazyTraversalPlanGenerator.h:87-120).
void LightPlan::Count1(const VList& um, const std::vector<bool>& nf, int level) {
  const int end = q_.VertexCount();
  if (level == kClNonMaterializeOpt && !nf[0]) {
    ops_.back().op = kComputeCount;
    return;
  }
  if (nf[0]) ops_.push_back({kFilterCompute, um[0]});
  ops_.push_back({kCount, end});
}

// GenerateCount2ExecuteSequence (LazyTraversalPlanGenerator.h:11-84).
void LightPlan::Count2(const VList& um, const std::vector<bool>& nf, int level) {
  const int end = q_.VertexCount();
  if (level == kClMaterialize) {
    ops_.push_back({kMaterialize, um[0]});
    ops_.push_back({kFilterCompute, um[1]});
    ops_.push_back({kCount, end});
    return;
  }
  if (level == kClNonMaterialize || (nf[0] && nf[1])) {
    if (nf[0]) ops_.push_back({kFilterCompute, um[0]});
    if (nf[1]) ops_.push_back({kFilterCompute, um[1]});
    ops_.push_back({kCount, end});
    return;
  }
  if (!nf[0] && !nf[1]) {
    // whichever of the two is touched last becomes COMPUTE_COUNT
    size_t first0 = ops_.size(), first1 = ops_.size();
    for (size_t j = 0; j < ops_.size(); ++j) {
      if (ops_[j].u == um[0] && first0 == ops_.size()) first0 = j;
      if (ops_[j].u == um[1] && first1 == ops_.size()) first1 = j;
    }
    ops_[std::max(first0, first1)].op = kComputeCount;
    return;
  }
  const int keep = nf[0] ? um[0] : um[1];
  const int fold = nf[0] ? um[1] : um[0];
  ops_.erase(ops_.begin() + OpIndex(kCompute, fold));
  ops_.push_back({kFilterCompute, keep});
  ops_.push_back({kComputeCount, fold});
}

// GenerateConstraints and the per-op condition sets
// (LazyTraversalPlanUtils.h:40-158).
void LightPlan::Constraints() {
  const int nv = q_.VertexCount();
  const size_t n = ops_.size();
  mat_.assign(n, VList());
  cun_.assign(n, VList());
  ccond_.assign(nv, CondList());
  mcond_.assign(nv, CondList());
  fcond_.assign(nv, CondList());
  VList materialized, computed;
  std::vector<bool> is_mat(nv, false);
  for (size_t i = 0; i < n; ++i) {
    mat_[i] = materialized;
    for (int u : computed)
      if (!is_mat[u]) cun_[i].push_back(u);
    const OpEntry& e = ops_[i];
    if (e.op == kCompute || e.op == kComputeCount || e.op == kComputePathCount) {
      // compute condition: against everything materialized so far
      CondList& cc = ccond_[e.u];
      cc.clear();
      for (int w : materialized) cc.push_back({q_.ConditionType(e.u, w), static_cast<uint8_t>(w)});
    }
    if (e.op == kMaterialize || e.op == kFilterCompute) {
      // against vertices materialized since this vertex was computed, latest first
      CondList& dst = (e.op == kMaterialize) ? mcond_[e.u] : fcond_[e.u];
      dst.clear();
      for (size_t j = i; j-- > 0;) {
        if (ops_[j].op == kCompute && ops_[j].u == e.u) break;
        if (ops_[j].op == kMaterialize)
          dst.push_back({q_.ConditionType(e.u, ops_[j].u), static_cast<uint8_t>(ops_[j].u)});
      }
    }
    if (e.op == kCompute || e.op == kComputePathCount) computed.push_back(e.u);
    if (e.op == kMaterialize) {
      materialized.push_back(e.u);
      is_mat[e.u] = true;
    }
  }
}

bool LightPlan::Build(const Query& q, int compress_level) {
  if (compress_level < kClMaterialize || compress_level > kClSpecial) return false;
  q_ = q;
  cl_ = compress_level;
  seq_ = LightOrders()[q.Type()];
  ops_.clear();
  ComputeSequence();
  VList um;
  std::vector<bool> nf;
  InspectUnmaterialized(&um, &nf);
  const int qt = q.Type();
  if (cl_ == kClSpecial && (qt == 2 || qt == 5 || qt == 6)) {
    // SpecialExecuteSequence (LazyTraversalPlanGenerator.h:244-300)
    if (qt == 5) {
      for (size_t i = 0; i < um.size(); ++i)
        if (nf[i]) ops_.push_back({kFilterCompute, um[i]});
    } else {
      ops_[ops_.size() - 2].op = kComputePathCount;
      ops_[ops_.size() - 1].op = kComputePathCount;
    }
    ops_.push_back({kCount, q.VertexCount()});
  } else {
    const int level = (cl_ == kClSpecial) ? kClNonMaterializeOpt : cl_;
    if (um.size() == 1) {
      Count1(um, nf, level);
    } else if (um.size() == 2) {
      Count2(um, nf, level);
    } else {
      for (size_t i = 2; i < um.size(); ++i) ops_.push_back({kMaterialize, um[i]});
      Count2(VList(um.begin(), um.begin() + 2), std::vector<bool>(2, true), level);
    }
  }
  Constraints();
  return true;
}

std::string LightPlan::Describe() const {
  std::string s = "Q" + std::to_string(q_.Type()) + " cl=" + std::to_string(cl_) +
                  " ord=" + (q_.OrderingEnabled() ? "1" : "0") + " seq=";
  for (size_t i = 0; i < seq_.size(); ++i) s += (i ? "," : "") + std::to_string(seq_[i]);
  s += "\nops:";
  for (const OpEntry& e : ops_) s += std::string(" ") + OpShort(e.op) + std::to_string(e.u);
  s += "\n";
  for (size_t i = 0; i < ops_.size(); ++i) {
    s += "L" + std::to_string(i) + " mat=";
    AppendV(&s, mat_[i]);
    s += " cun=";
    AppendV(&s, cun_[i]);
    s += "\n";
  }
  for (int u = 0; u < q_.VertexCount(); ++u) {
    s += "V" + std::to_string(u) + " bconn=";
    AppendV(&s, bconn_[u]);
    s += " ccond=";
    AppendC(&s, ccond_[u]);
    s += " mcond=";
    AppendC(&s, mcond_[u]);
    s += " fcond=";
    AppendC(&s, fcond_[u]);
    s += "\n";
  }
  return s;
}

bool CanDirectCount(int qt) { return qt >= 0 && qt <= 11 && qt != 4 && qt != 9; }

namespace {

void FillList(uint8_t* dst, uint8_t* n, const VList& v) {
  *n = static_cast<uint8_t>(v.size());
  for (size_t i = 0; i < v.size(); ++i) dst[i] = static_cast<uint8_t>(v[i]);
}

void FillConds(Cond* dst, uint8_t* n, const CondList& c) {
  *n = static_cast<uint8_t>(c.size());
  for (size_t i = 0; i < c.size(); ++i) dst[i] = c[i];
}

// Every pair of query vertices adjacent.
bool IsClique(const Query& q) {
  for (int u = 0; u < q.VertexCount(); ++u)
    if (static_cast<int>(q.Connectivity()[u].size()) != q.VertexCount() - 1) return false;
  return true;
}

}  // namespace

bool MakeLightEnginePlan(int query_type, int compress_level, bool enable_ordering,
                         EnginePlan* out, std::string* describe) {
  Query q;
  if (!q.Init(query_type, enable_ordering)) return false;
  LightPlan p;
  if (!p.Build(q, compress_level)) return false;
  const auto& ops = p.ExecuteOperations();
  if (ops.size() > static_cast<size_t>(kMaxOps)) return false;
  *out = EnginePlan{};
  out->algo = kAlgoLight;
  out->query_type = query_type;
  out->compress_level = compress_level;
  out->enable_ordering = enable_ordering ? 1 : 0;
  out->nv = q.VertexCount();
  out->nops = static_cast<int>(ops.size());
  out->direct_count = (compress_level == kClSpecial && CanDirectCount(query_type)) ? 1 : 0;
  out->reuse = 1;
  out->clique = IsClique(q) ? 1 : 0;
  for (size_t i = 0; i < ops.size(); ++i) {
    OpDesc& d = out->ops[i];
    d.kind = ops[i].op;
    d.u = static_cast<uint8_t>(ops[i].u);
    FillList(d.mat, &d.nmat, p.MaterializedVertices()[i]);
    FillList(d.cun, &d.ncun, p.ComputedUnmaterializedVertices()[i]);
    d.pair_op = kNonEqual;
    switch (ops[i].op) {
      case kCompute:
      case kComputePathCount:
      case kComputeCount:
        FillList(d.bconn, &d.nb, p.BackwardConnectivity()[ops[i].u]);
        FillConds(d.cond, &d.nc, p.ComputeCondition()[ops[i].u]);
        if (ops[i].op == kComputeCount && d.ncun == 1)
          d.pair_op = q.ConditionType(ops[i].u, d.cun[0]);
        break;
      case kMaterialize:
        FillConds(d.cond, &d.nc, p.MaterializeCondition()[ops[i].u]);
        break;
      case kFilterCompute:
        FillConds(d.cond, &d.nc, p.FilterCondition()[ops[i].u]);
        break;
      case kCount:
        if (d.ncun == 2) d.pair_op = q.ConditionType(d.cun[0], d.cun[1]);
        break;
    }
  }
  if (describe) *describe = p.Describe();
  return true;
}

bool MakeGpsmEnginePlan(int query_type, bool enable_ordering, bool reuse, EnginePlan* out,
                        std::string* describe) {
  Query q;
  if (!q.Init(query_type, enable_ordering)) return false;
  const int nv = q.VertexCount();
  if (2 * nv > kMaxOps) return false;
  *out = EnginePlan{};
  out->algo = reuse ? kAlgoGpsmReuse : kAlgoGpsm;
  out->reuse = reuse ? 1 : 0;
  out->query_type = query_type;
  out->compress_level = -1;
  out->enable_ordering = enable_ordering ? 1 : 0;
  out->nv = nv;
  out->clique = IsClique(q) ? 1 : 0;
  std::string text = "GPSM Q" + std::to_string(query_type) + " ord=" +
                     (enable_ordering ? "1" : "0") + "\n";
  int k = 0;
  VList materialized;
  for (int u = 0; u < nv; ++u) {
    // GetOrderedConnectivity (Plan.h:60-71) and GetWholeOrderedOrdering
    // (Plan.h:163-174) under the identity match order.
    VList back;
    for (int w : q.Connectivity()[u])
      if (w < u) back.push_back(w);
    CondList whole;
    for (int w = 0; w < u; ++w) whole.push_back({q.ConditionType(u, w), static_cast<uint8_t>(w)});
    if (u > 0) {
      if (back.empty()) return false;
      OpDesc& c = out->ops[k++];
      c.kind = kCompute;
      c.u = static_cast<uint8_t>(u);
      FillList(c.bconn, &c.nb, back);
      FillConds(c.cond, &c.nc, whole);
      FillList(c.mat, &c.nmat, materialized);
      c.ncun = 0;
      c.pair_op = kNonEqual;
      text += "L" + std::to_string(u) + " bconn=";
      AppendV(&text, back);
      text += " cond=";
      AppendC(&text, whole);
      text += "\n";
    }
    OpDesc& m = out->ops[k++];
    m.kind = kMaterialize;
    m.u = static_cast<uint8_t>(u);
    m.nc = 0;
    FillList(m.mat, &m.nmat, materialized);
    if (u > 0) {
      m.ncun = 1;
      m.cun[0] = static_cast<uint8_t>(u);
    }
    m.pair_op = kNonEqual;
    materialized.push_back(u);
  }
  out->nops = k;
  if (describe) *describe = text;
  return true;
}

}  // namespace sme
