// ref_plan_dump.cpp -- prints the reference's own LIGHT plans in the oracle's
// canonical text form (or_plan_text in oracle.c).
//
// TEST INFRASTRUCTURE ONLY.  Built by `make -C oracle ref` against the
// reference headers where they lie (REF_SRC); output goes to oracle/_ref/.
// It only calls the reference's plan classes:
//   LazyTraversalPlan (src/query/LazyTraversalPlan.h:13-29, 173-197)
//   Plan::GetBackwardConnectivity (src/query/common/Plan.h:202-217)
//   GetComputeCondition / GetMaterializeCondition / GetFilterCondition
//     (src/query/LazyTraversalPlan.h:329-349)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "LazyTraversalPlan.h"

static const char* OpName(LazyTraversalOperation op) {
  switch (op) {
    case COMPUTE: return "C";
    case COMPUTE_PATH_COUNT: return "CPC";
    case MATERIALIZE: return "M";
    case FILTER_COMPUTE: return "FC";
    case COUNT: return "COUNT";
    case COMPUTE_COUNT: return "CC";
  }
  return "?";
}

static void PutV(const std::vector<uintV>& v) {
  std::printf("[");
  for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ",%u" : "%u", v[i]);
  std::printf("]");
}

static void PutC(const std::vector<CondType>& c) {
  std::printf("[");
  for (size_t i = 0; i < c.size(); ++i) {
    char k = c[i].first == LESS_THAN ? '<' : c[i].first == LARGER_THAN ? '>' : c[i].first == NON_EQUAL ? '!' : '?';
    std::printf(i ? ",%c%u" : "%c%u", k, c[i].second);
  }
  std::printf("]");
}

int main(int argc, char** argv) {
  for (int ord = 1; ord >= 0; --ord) {
    for (int q = 0; q < 14; ++q) {
      for (int cl = 0; cl < 4; ++cl) {
        Query query((QueryType)q, ord != 0);
        LazyTraversalPlan plan(&query, 1, 1, IN_MEMORY_GPU, O0, (LazyTraversalCompressLevel)cl);
        size_t nv = plan.GetVertexCount();
        SearchSequence& seq = plan.GetSearchSequence();
        std::printf("Q%d cl=%d ord=%d seq=", q, cl, ord);
        for (size_t i = 0; i < nv; ++i) std::printf(i ? ",%u" : "%u", seq[i]);
        std::printf("\nops:");
        auto& ops = plan.GetExecuteOperations();
        for (auto& e : ops) std::printf(" %s%u", OpName(e.first), e.second);
        std::printf("\n");
        auto& mat = plan.GetMaterializedVertices();
        auto& cun = plan.GetComputedUnmaterializedVertices();
        for (size_t i = 0; i < ops.size(); ++i) {
          std::printf("L%zu mat=", i);
          PutV(mat[i]);
          std::printf(" cun=");
          PutV(cun[i]);
          std::printf("\n");
        }
        AllConnType bconn;
        Plan::GetBackwardConnectivity(bconn, plan.GetConnectivity(), seq);
        AllCondType cc, mc, fc;
        plan.GetComputeCondition(cc);
        plan.GetMaterializeCondition(mc);
        plan.GetFilterCondition(fc);
        for (size_t u = 0; u < nv; ++u) {
          std::printf("V%zu bconn=", u);
          PutV(bconn[u]);
          std::printf(" ccond=");
          PutC(cc[u]);
          std::printf(" mcond=");
          PutC(mc[u]);
          std::printf(" fcond=");
          PutC(fc[u]);
          std::printf("\n");
        }
      }
    }
  }
  return 0;
}
