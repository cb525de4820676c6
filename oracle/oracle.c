/*
 * oracle.c -- CPU restatement of the reference's subgraph-enumeration path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Never linked into the product.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_query_init / or_light_plan : the canonical plan text of all 14 queries
 *       x 4 compress levels equals the text printed by the reference's own
 *       plan classes (oracle/_ref/plan_dump, fixture tests/golden/light_plans.txt).
 *   or_wcoj_count : counts equal the reference CPU binary
 *       (oracle/_ref/patternmatchcpu -a 0) on seeded RMAT/ER graphs (fixture
 *       tests/golden/ref_counts.json), the K_n closed forms, and a networkx
 *       brute-force monomorphism count / |Aut(P)| on tiny random graphs.
 *   or_light_count : total equals or_wcoj_count on every fixture graph and
 *       query (full start range); per-level outputs are pinned by the closed
 *       forms level(M1)=level(C1)=nnz/2 for the triangle and by the
 *       invariants checked in tests/test_oracle.py (Q0: {n, nnz/2, nnz/2, T}).
 *   or_gpsm_count : total equals or_wcoj_count; per-level path counts equal
 *       the reference's GPSM definition (size of each BFS level) and, for the
 *       last level, the final count.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

/* ------------------------------------------------------------------------ */
/* Query registry: query/common/Query.h:84-433.  Adjacency lists are given in
 * the push_back order of the reference (order matters: it fixes the order of
 * backward neighbours and therefore the generated execute sequence).  Orders
 * are "<k" (LESS_THAN k) and ">k" (LARGER_THAN k), also in push order. */
typedef struct { int nv; const char* con[OR_MAXQ]; const char* ord[OR_MAXQ]; } q_desc;

static const q_desc kQueries[14] = {
  /* Q0 triangle            Query.h:84-100 */
  {3, {"12", "02", "01"}, {"<1", ">0<2", ">1"}},
  /* Q1 square              Query.h:101-123 */
  {4, {"13", "02", "13", "02"}, {"<1<2<3", ">0<3", ">0", ">0>1"}},
  /* Q2 chordal square      Query.h:124-144 */
  {4, {"123", "02", "013", "02"}, {"<2", "<3", ">0", ">1"}},
  /* Q3 4-clique            Query.h:145-169 */
  {4, {"123", "023", "013", "012"}, {"<1", ">0<2", ">1<3", ">2"}},
  /* Q4 house               Query.h:170-190 */
  {5, {"12", "023", "014", "14", "23"}, {"", "<2", ">1", "", ""}},
  /* Q5 quad triangle       Query.h:192-218 */
  {6, {"12345", "02", "013", "024", "035", "04"}, {"", "", "<4", "", ">2", ""}},
  /* Q6 near-5-clique       Query.h:220-246 */
  {5, {"1234", "023", "013", "0124", "03"}, {"<3", "<2", ">1", ">0", ""}},
  /* Q7 5-clique            Query.h:248-282 */
  {5, {"1234", "0234", "0134", "0124", "0123"}, {"<1", ">0<2", ">1<3", ">2<4", ">3"}},
  /* Q8 chordal roof        Query.h:283-307 */
  {5, {"123", "023", "014", "014", "23"}, {"<1", ">0", "<3", ">2", ""}},
  /* Q9 three triangle      Query.h:308-325 */
  {5, {"1234", "032", "014", "01", "02"}, {"", "<2", ">1", "", ""}},
  /* Q10 solar square       Query.h:326-356 */
  {5, {"1234", "024", "013", "024", "013"}, {"", "<2<3<4", "<4>1", ">1", ">1>2"}},
  /* Q11 6-clique           Query.h:358-373 */
  {6, {"12345", "02345", "01345", "01245", "01235", "01234"},
   {"<1", ">0<2", ">1<3", ">2<4", ">3<5", ">4"}},
  /* Q12 near-near-5-clique Query.h:375-403 */
  {5, {"1234", "0234", "013", "0124", "013"}, {"<1", ">0", "<4", "", ">2"}},
  /* Q13 triangle in diamond Query.h:405-433 */
  {6, {"1235", "0234", "0145", "01", "12", "02"}, {"<1", "<2>0", ">1", "", "", ""}},
};

/* LazyTraversalPlanHardCode.h:59-117 */
static const int kLightSeq[14][OR_MAXQ] = {
  {0, 1, 2}, {0, 1, 2, 3}, {0, 2, 1, 3}, {0, 1, 2, 3}, {1, 2, 3, 0, 4},
  {0, 2, 4, 1, 3, 5}, {0, 1, 3, 2, 4}, {0, 1, 2, 3, 4}, {0, 1, 2, 3, 4},
  {0, 1, 2, 3, 4}, {0, 1, 3, 2, 4}, {0, 1, 2, 3, 4, 5}, {0, 1, 3, 2, 4},
  {0, 1, 2, 3, 4, 5}};

int or_query_init(int q, int enable_ordering, or_query* out) {
  if (q < 0 || q >= 14 || !out) return -1;
  memset(out, 0, sizeof(*out));
  const q_desc* d = &kQueries[q];
  out->q = q;
  out->nv = d->nv;
  for (int u = 0; u < d->nv; ++u) {
    for (const char* s = d->con[u]; *s; ++s) out->con[u].v[out->con[u].cnt++] = *s - '0';
    if (!enable_ordering) continue; /* Query.h:77-82 clears the orders */
    for (const char* s = d->ord[u]; *s; s += 2) {
      int c = out->order[u].cnt++;
      out->order[u].op[c] = (s[0] == '<') ? OR_LT : OR_GT;
      out->order[u].v[c] = s[1] - '0';
    }
  }
  return 0;
}

/* Plan::GetConditionType, Plan.h:32-43: NON_EQUAL unless an order is given. */
static int cond_type(const or_cset* order, int l1, int l2) {
  for (int i = 0; i < order[l1].cnt; ++i)
    if (order[l1].v[i] == l2) return order[l1].op[i];
  return OR_NE;
}

static void vs_push(or_vset* s, int v) { s->v[s->cnt++] = v; }
static void cs_push(or_cset* s, int op, int v) { s->op[s->cnt] = op; s->v[s->cnt] = v; s->cnt++; }

static int op_index(const or_plan* p, int op, int u) {
  /* LazyTraversalPlanUtils::GetOperationIndex, LazyTraversalPlanUtils.h:9-18 */
  for (int i = 0; i < p->nops; ++i)
    if (p->op[i] == op && p->opv[i] == u) return i;
  return p->nops;
}

static void push_op(or_plan* p, int op, int u) { p->op[p->nops] = op; p->opv[p->nops] = u; p->nops++; }

static void erase_op(or_plan* p, int idx) {
  for (int i = idx; i + 1 < p->nops; ++i) { p->op[i] = p->op[i + 1]; p->opv[i] = p->opv[i + 1]; }
  p->nops--;
}

/* LazyTraversalPlanGenerator::GenerateCount2ExecuteSequence (:11-84) */
static void gen_count2(or_plan* p, const int* um, const int* nf, int cl) {
  int vc = p->nv;
  if (cl == OR_CL_MATERIALIZE) {
    push_op(p, OR_MATERIALIZE, um[0]);
    push_op(p, OR_FILTER_COMPUTE, um[1]);
    push_op(p, OR_COUNT, vc);
  } else if (cl == OR_CL_NON_MATERIALIZE) {
    if (nf[0]) push_op(p, OR_FILTER_COMPUTE, um[0]);
    if (nf[1]) push_op(p, OR_FILTER_COMPUTE, um[1]);
    push_op(p, OR_COUNT, vc);
  } else {
    if (nf[0] && nf[1]) {
      push_op(p, OR_FILTER_COMPUTE, um[0]);
      push_op(p, OR_FILTER_COMPUTE, um[1]);
      push_op(p, OR_COUNT, vc);
    } else if (!nf[0] && !nf[1]) {
      /* the later of the two last operations becomes COMPUTE_COUNT */
      int last0 = p->nops, last1 = p->nops;
      for (int j = p->nops - 1; j >= 0; --j) {
        if (p->opv[j] == um[0]) last0 = j;
        if (p->opv[j] == um[1]) last1 = j;
      }
      int mx = last0 > last1 ? last0 : last1;
      p->op[mx] = OR_COMPUTE_COUNT;
    } else {
      int need = nf[0] ? um[0] : um[1];
      int non_need = nf[0] ? um[1] : um[0];
      erase_op(p, op_index(p, OR_COMPUTE, non_need));
      push_op(p, OR_FILTER_COMPUTE, need);
      push_op(p, OR_COMPUTE_COUNT, non_need);
    }
  }
}

/* LazyTraversalPlanGenerator::GenerateCount1ExecuteSequence (:87-120) */
static void gen_count1(or_plan* p, const int* um, const int* nf, int cl) {
  int vc = p->nv;
  if (cl == OR_CL_MATERIALIZE || cl == OR_CL_NON_MATERIALIZE) {
    if (nf[0]) push_op(p, OR_FILTER_COMPUTE, um[0]);
    push_op(p, OR_COUNT, vc);
  } else {
    if (nf[0]) {
      push_op(p, OR_FILTER_COMPUTE, um[0]);
      push_op(p, OR_COUNT, vc);
    } else {
      p->op[p->nops - 1] = OR_COMPUTE_COUNT;
    }
  }
}

int or_light_plan(int q, int cl, int enable_ordering, or_plan* p) {
  or_query qq;
  if (or_query_init(q, enable_ordering, &qq)) return -1;
  if (cl < 0 || cl > OR_CL_SPECIAL) return -1;
  memset(p, 0, sizeof(*p));
  p->q = q;
  p->compress_level = cl;
  p->enable_ordering = enable_ordering;
  p->nv = qq.nv;
  memcpy(p->seq, kLightSeq[q], sizeof(int) * OR_MAXQ);
  for (int u = 0; u < qq.nv; ++u) p->order[u] = qq.order[u];

  /* Plan::GetBackwardConnectivity, Plan.h:202-217 (original vertex ids) */
  {
    int vis[OR_MAXQ] = {0};
    for (int i = 0; i < qq.nv; ++i) {
      int u = p->seq[i];
      for (int j = 0; j < qq.con[u].cnt; ++j)
        if (vis[qq.con[u].v[j]]) vs_push(&p->bconn[u], qq.con[u].v[j]);
      vis[u] = 1;
    }
  }
  /* GenerateComputeExecuteSequence, LazyTraversalPlanGenerator.h:124-154
   * (materialize_first_edge = false for the intra-partition plan) */
  {
    int mat[OR_MAXQ] = {0};
    push_op(p, OR_MATERIALIZE, p->seq[0]);
    mat[p->seq[0]] = 1;
    for (int i = 1; i < qq.nv; ++i) {
      int u = p->seq[i];
      for (int j = 0; j < p->bconn[u].cnt; ++j) {
        int pu = p->bconn[u].v[j];
        if (!mat[pu]) { push_op(p, OR_MATERIALIZE, pu); mat[pu] = 1; }
      }
      push_op(p, OR_COMPUTE, u);
    }
  }
  /* InspectUnmaterializedVertices, LazyTraversalPlanGenerator.h:158-184 */
  int um[OR_MAXQ], nf[OR_MAXQ], num = 0;
  {
    int mat[OR_MAXQ] = {0};
    for (int i = 0; i < p->nops; ++i) if (p->op[i] == OR_MATERIALIZE) mat[p->opv[i]] = 1;
    for (int u = 0; u < qq.nv; ++u) {
      if (mat[u]) continue;
      int ind = op_index(p, OR_COMPUTE, u), f = 0;
      for (int i = ind + 1; i < p->nops; ++i) if (p->op[i] == OR_MATERIALIZE) f = 1;
      um[num] = u; nf[num] = f; num++;
    }
  }
  /* LazyTraversalPlan::GenConstraints, LazyTraversalPlan.h:173-197 */
  if (cl == OR_CL_SPECIAL && (q == 2 || q == 5 || q == 6)) {
    /* SpecialExecuteSequence, LazyTraversalPlanGenerator.h:244-300 */
    if (q == 2 || q == 6) {
      p->op[p->nops - 2] = OR_COMPUTE_PATH_COUNT;
      p->op[p->nops - 1] = OR_COMPUTE_PATH_COUNT;
      push_op(p, OR_COUNT, qq.nv);
    } else {
      for (int i = 0; i < num; ++i) if (nf[i]) push_op(p, OR_FILTER_COMPUTE, um[i]);
      push_op(p, OR_COUNT, qq.nv);
    }
  } else {
    int lvl = (cl == OR_CL_SPECIAL) ? OR_CL_NON_MATERIALIZE_OPT : cl;
    if (num == 1) {
      gen_count1(p, um, nf, lvl);
    } else if (num == 2) {
      gen_count2(p, um, nf, lvl);
    } else {
      for (int i = 2; i < num; ++i) push_op(p, OR_MATERIALIZE, um[i]);
      int nf2[2] = {1, 1};
      gen_count2(p, um, nf2, lvl);
    }
  }
  /* LazyTraversalPlanUtils::GenerateConstraints (:40-68) */
  {
    int mat[OR_MAXQ] = {0};
    or_vset all_mat = {0}, all_comp = {0};
    for (int i = 0; i < p->nops; ++i) {
      p->mat[i] = all_mat;
      for (int j = 0; j < all_comp.cnt; ++j)
        if (!mat[all_comp.v[j]]) vs_push(&p->cun[i], all_comp.v[j]);
      if (p->op[i] == OR_COMPUTE || p->op[i] == OR_COMPUTE_PATH_COUNT) {
        vs_push(&all_comp, p->opv[i]);
      } else if (p->op[i] == OR_MATERIALIZE) {
        vs_push(&all_mat, p->opv[i]);
        mat[p->opv[i]] = 1;
      }
    }
  }
  /* GetComputeCondition (:73-92) */
  {
    or_vset m = {0};
    for (int i = 0; i < p->nops; ++i) {
      int op = p->op[i];
      if (op == OR_COMPUTE || op == OR_COMPUTE_COUNT || op == OR_COMPUTE_PATH_COUNT) {
        int u = p->opv[i];
        p->ccond[u].cnt = 0;
        for (int j = 0; j < m.cnt; ++j) cs_push(&p->ccond[u], cond_type(p->order, u, m.v[j]), m.v[j]);
      } else if (op == OR_MATERIALIZE) {
        vs_push(&m, p->opv[i]);
      }
    }
  }
  /* GetMaterializeCondition (:97-125) and GetFilterCondition (:131-158) */
  for (int i = 0; i < p->nops; ++i) {
    int op = p->op[i];
    if (op != OR_MATERIALIZE && op != OR_FILTER_COMPUTE) continue;
    int u = p->opv[i];
    or_cset* dst = (op == OR_MATERIALIZE) ? &p->mcond[u] : &p->fcond[u];
    dst->cnt = 0;
    for (int j = i - 1; j >= 0; --j) {
      if (p->op[j] == OR_COMPUTE && p->opv[j] == u) break;
      if (p->op[j] == OR_MATERIALIZE) cs_push(dst, cond_type(p->order, u, p->opv[j]), p->opv[j]);
    }
  }
  return 0;
}

static const char* kOpName[] = {"C", "CPC", "M", "FC", "COUNT", "CC"};

static int put_vset(char* buf, size_t cap, size_t* pos, const or_vset* s) {
  int w = snprintf(buf + *pos, cap > *pos ? cap - *pos : 0, "[");
  *pos += w;
  for (int i = 0; i < s->cnt; ++i) {
    w = snprintf(buf + *pos, cap > *pos ? cap - *pos : 0, i ? ",%d" : "%d", s->v[i]);
    *pos += w;
  }
  w = snprintf(buf + *pos, cap > *pos ? cap - *pos : 0, "]");
  *pos += w;
  return 0;
}

static int put_cset(char* buf, size_t cap, size_t* pos, const or_cset* s) {
  static const char kc[] = {'<', '>', '!', '?'};
  int w = snprintf(buf + *pos, cap > *pos ? cap - *pos : 0, "[");
  *pos += w;
  for (int i = 0; i < s->cnt; ++i) {
    w = snprintf(buf + *pos, cap > *pos ? cap - *pos : 0, i ? ",%c%d" : "%c%d", kc[s->op[i]], s->v[i]);
    *pos += w;
  }
  w = snprintf(buf + *pos, cap > *pos ? cap - *pos : 0, "]");
  *pos += w;
  return 0;
}

#define PUTS(...) do { int w_ = snprintf(buf + pos, cap > pos ? cap - pos : 0, __VA_ARGS__); pos += w_; } while (0)

int or_plan_text(const or_plan* p, char* buf, size_t cap) {
  size_t pos = 0;
  PUTS("Q%d cl=%d ord=%d seq=", p->q, p->compress_level, p->enable_ordering);
  for (int i = 0; i < p->nv; ++i) PUTS(i ? ",%d" : "%d", p->seq[i]);
  PUTS("\nops:");
  for (int i = 0; i < p->nops; ++i) PUTS(" %s%d", kOpName[p->op[i]], p->opv[i]);
  PUTS("\n");
  for (int i = 0; i < p->nops; ++i) {
    PUTS("L%d mat=", i); put_vset(buf, cap, &pos, &p->mat[i]);
    PUTS(" cun="); put_vset(buf, cap, &pos, &p->cun[i]);
    PUTS("\n");
  }
  for (int u = 0; u < p->nv; ++u) {
    PUTS("V%d bconn=", u); put_vset(buf, cap, &pos, &p->bconn[u]);
    PUTS(" ccond="); put_cset(buf, cap, &pos, &p->ccond[u]);
    PUTS(" mcond="); put_cset(buf, cap, &pos, &p->mcond[u]);
    PUTS(" fcond="); put_cset(buf, cap, &pos, &p->fcond[u]);
    PUTS("\n");
  }
  return (int)pos;
}

/* ------------------------------------------------------------------------ */
/* Shared graph helpers.                                                     */

typedef struct {
  uint64_t n;
  const uint64_t* rp;
  const uint32_t* cols;
} graph_t;

static inline int cond_ok(int op, uint32_t x, uint32_t y) {
  /* GpuUtils::Filter::CheckVertexCondition, GPUFilter.cuh:10-23;
   * CheckCondition, CPUFilter.h:7-24 */
  switch (op) {
    case OR_LT: return x < y;
    case OR_GT: return x > y;
    case OR_NE: return x != y;
    default: return 1;
  }
}

/* Sorted merge intersection, CPUIntersection.h:13-33. */
static size_t merge_isect(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out) {
  size_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else { out[k++] = a[i]; ++i; ++j; }
  }
  return k;
}

/* MWayIntersect, CPUIntersection.h:64-99: successive pairwise merges of the
 * adjacency lists of path[levels[0..k)]. `tmp` must hold max degree. */
static size_t mway(const graph_t* g, const uint32_t* path, const int* levels, int k, uint32_t* out, uint32_t* tmp) {
  uint32_t v = path[levels[0]];
  size_t cnt = g->rp[v + 1] - g->rp[v];
  memcpy(out, g->cols + g->rp[v], cnt * sizeof(uint32_t));
  for (int j = 1; j < k; ++j) {
    uint32_t w = path[levels[j]];
    size_t c2 = merge_isect(g->cols + g->rp[w], g->rp[w + 1] - g->rp[w], out, cnt, tmp);
    memcpy(out, tmp, c2 * sizeof(uint32_t));
    cnt = c2;
  }
  return cnt;
}

static uint64_t max_degree(const graph_t* g) {
  uint64_t m = 0;
  for (uint64_t u = 0; u < g->n; ++u) {
    uint64_t d = g->rp[u + 1] - g->rp[u];
    if (d > m) m = d;
  }
  return m;
}

/* ------------------------------------------------------------------------ */
/* CPU_WCOJ: CPUWCOJ.h:20-111 with TraversalPlan's identity order
 * (TraversalPlan.h:193-233), GetOrderedConnectivity (Plan.h:60-71) and
 * GetOrderedOrdering (Plan.h:77-88). */
typedef struct {
  const graph_t* g;
  int nv;
  int lv[OR_MAXQ][OR_MAXQ];
  int nlv[OR_MAXQ];
  or_cset cond[OR_MAXQ];
  uint32_t* buf[OR_MAXQ];
  uint32_t* tmp;
} wcoj_ctx;

static uint64_t wcoj_dfs(wcoj_ctx* c, int lvl, uint32_t* path) {
  if (lvl >= c->nv || lvl >= OR_MAXQ) return 1;
  uint64_t ans = 0;
  const or_cset* cd = &c->cond[lvl];
  if (c->nlv[lvl] == 0) {
    for (uint64_t x = 0; x < c->g->n; ++x) {
      int ok = 1;
      for (int j = 0; j < cd->cnt && ok; ++j) ok = cond_ok(cd->op[j], (uint32_t)x, path[cd->v[j]]);
      for (int j = 0; j < lvl && ok; ++j) ok = path[j] != x; /* CheckEquality */
      if (!ok) continue;
      path[lvl] = (uint32_t)x;
      ans += wcoj_dfs(c, lvl + 1, path);
    }
    return ans;
  }
  size_t k = mway(c->g, path, c->lv[lvl], c->nlv[lvl], c->buf[lvl], c->tmp);
  uint32_t* cand = c->buf[lvl];
  for (size_t i = 0; i < k; ++i) {
    uint32_t x = cand[i];
    int ok = 1;
    for (int j = 0; j < cd->cnt && ok; ++j) ok = cond_ok(cd->op[j], x, path[cd->v[j]]);
    for (int j = 0; j < lvl && ok; ++j) ok = path[j] != x;
    if (!ok) continue;
    path[lvl] = x;
    ans += wcoj_dfs(c, lvl + 1, path);
  }
  return ans;
}

static int wcoj_run(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                    int enable_ordering, uint32_t v_begin, uint32_t v_end, const uint32_t* starts,
                    int nthreads, uint64_t* out_count);

int or_wcoj_count(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                  int enable_ordering, uint32_t v_begin, uint32_t v_end, int nthreads,
                  uint64_t* out_count) {
  if (v_end > n) v_end = (uint32_t)n;
  if (v_begin > v_end) v_begin = v_end;
  return wcoj_run(n, row_ptrs, cols, q, enable_ordering, v_begin, v_end, NULL, nthreads, out_count);
}

int or_wcoj_count_list(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                       int enable_ordering, const uint32_t* starts, uint64_t nstarts,
                       int nthreads, uint64_t* out_count) {
  for (uint64_t i = 0; i < nstarts; ++i)
    if (starts[i] >= n) return -1;
  return wcoj_run(n, row_ptrs, cols, q, enable_ordering, 0, (uint32_t)nstarts, starts, nthreads,
                  out_count);
}

/* starts == NULL: start vertices v_begin..v_end-1; else starts[v_begin..v_end). */
static int wcoj_run(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                    int enable_ordering, uint32_t v_begin, uint32_t v_end, const uint32_t* starts,
                    int nthreads, uint64_t* out_count) {
  or_query qq;
  if (or_query_init(q, enable_ordering, &qq)) return -1;
  graph_t g = {n, row_ptrs, cols};
  uint64_t md = max_degree(&g) + 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  uint64_t total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
  {
    wcoj_ctx c;
    memset(&c, 0, sizeof(c));
    c.g = &g;
    c.nv = qq.nv;
    for (int u = 0; u < qq.nv; ++u) {
      for (int j = 0; j < qq.con[u].cnt; ++j)
        if (qq.con[u].v[j] < u) c.lv[u][c.nlv[u]++] = qq.con[u].v[j];
      for (int j = 0; j < qq.order[u].cnt; ++j)
        if (qq.order[u].v[j] < u) cs_push(&c.cond[u], qq.order[u].op[j], qq.order[u].v[j]);
      c.buf[u] = (uint32_t*)malloc(md * sizeof(uint32_t));
    }
    c.tmp = (uint32_t*)malloc(md * sizeof(uint32_t));
    uint32_t path[OR_MAXQ];
#pragma omp for schedule(dynamic, 64)
    for (int64_t v = v_begin; v < (int64_t)v_end; ++v) {
      path[0] = starts ? starts[v] : (uint32_t)v;
      total += wcoj_dfs(&c, 1, path);
    }
    for (int u = 0; u < qq.nv; ++u) free(c.buf[u]);
    free(c.tmp);
  }
  *out_count = total;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* GPSM: HybridGPSMGPUProcessor.cuh:23-33 (count = size of the last level),
 * GPSMCheckConstraints.cuh:40-105 (intersect GetOrderedConnectivity lists,
 * filter with GetWholeOrderedOrdering, Plan.h:163-174). */
typedef struct {
  const graph_t* g;
  int nv;
  int lv[OR_MAXQ][OR_MAXQ];
  int nlv[OR_MAXQ];
  or_cset cond[OR_MAXQ];
  uint32_t* buf[OR_MAXQ];
  uint32_t* tmp;
  uint64_t level[OR_MAXQ];
  /* match-list output (sequential walk only) */
  uint32_t* out;
  uint64_t cap, rows;
} gpsm_ctx;

static void gpsm_dfs(gpsm_ctx* c, int lvl, uint32_t* path) {
  if (lvl >= c->nv || lvl >= OR_MAXQ) return;
  size_t k = mway(c->g, path, c->lv[lvl], c->nlv[lvl], c->buf[lvl], c->tmp);
  const or_cset* cd = &c->cond[lvl];
  uint32_t* cand = c->buf[lvl];
  for (size_t i = 0; i < k; ++i) {
    uint32_t x = cand[i];
    int ok = 1;
    for (int j = 0; j < cd->cnt && ok; ++j) ok = cond_ok(cd->op[j], x, path[cd->v[j]]);
    if (!ok) continue;
    path[lvl] = x;
    c->level[lvl]++;
    if (lvl == c->nv - 1 && c->out && c->rows < c->cap) {
      memcpy(c->out + c->rows * (uint64_t)c->nv, path, sizeof(uint32_t) * (size_t)c->nv);
      c->rows++;
    }
    gpsm_dfs(c, lvl + 1, path);
  }
}

int or_gpsm_count(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                  int enable_ordering, uint32_t v_begin, uint32_t v_end, int nthreads,
                  uint64_t* out_count, uint64_t* level_out) {
  or_query qq;
  if (or_query_init(q, enable_ordering, &qq)) return -1;
  if (v_end > n) v_end = (uint32_t)n;
  graph_t g = {n, row_ptrs, cols};
  uint64_t md = max_degree(&g) + 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  uint64_t lv_tot[OR_MAXQ] = {0};
  for (int u = 1; u < qq.nv; ++u) {
    int has = 0;
    for (int j = 0; j < qq.con[u].cnt; ++j) has |= qq.con[u].v[j] < u;
    if (!has) return -2; /* GPSM needs a backward neighbour at every level */
  }
#pragma omp parallel num_threads(nthreads)
  {
    gpsm_ctx c;
    memset(&c, 0, sizeof(c));
    c.g = &g;
    c.nv = qq.nv;
    for (int u = 0; u < qq.nv; ++u) {
      for (int j = 0; j < qq.con[u].cnt; ++j)
        if (qq.con[u].v[j] < u) c.lv[u][c.nlv[u]++] = qq.con[u].v[j];
      for (int v = 0; v < u; ++v) cs_push(&c.cond[u], cond_type(qq.order, u, v), v);
      c.buf[u] = (uint32_t*)malloc(md * sizeof(uint32_t));
    }
    c.tmp = (uint32_t*)malloc(md * sizeof(uint32_t));
    uint32_t path[OR_MAXQ];
#pragma omp for schedule(dynamic, 64)
    for (int64_t v = v_begin; v < (int64_t)v_end; ++v) {
      path[0] = (uint32_t)v;
      c.level[0]++;
      gpsm_dfs(&c, 1, path);
    }
#pragma omp critical
    for (int u = 0; u < qq.nv; ++u) lv_tot[u] += c.level[u];
    for (int u = 0; u < qq.nv; ++u) free(c.buf[u]);
    free(c.tmp);
  }
  *out_count = lv_tot[qq.nv - 1];
  if (level_out) for (int u = 0; u < qq.nv; ++u) level_out[u] = lv_tot[u];
  return 0;
}

/* The GPSM matches themselves, in the order a sequential depth-first walk
 * over sorted candidate sets visits them: lexicographic in (M[0], M[1],
 * ...).  Pinned without re-typing: the row count equals or_gpsm_count, and
 * the tests check every row against the pattern's edges and conditions
 * directly and check that rows are distinct and sorted, which together fix
 * the list. */
int or_gpsm_list(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                 int enable_ordering, uint32_t v_begin, uint32_t v_end, uint32_t* out,
                 uint64_t cap, uint64_t* rows, uint64_t* total) {
  or_query qq;
  if (or_query_init(q, enable_ordering, &qq)) return -1;
  if (v_end > n) v_end = (uint32_t)n;
  for (int u = 1; u < qq.nv; ++u) {
    int has = 0;
    for (int j = 0; j < qq.con[u].cnt; ++j) has |= qq.con[u].v[j] < u;
    if (!has) return -2;
  }
  graph_t g = {n, row_ptrs, cols};
  uint64_t md = max_degree(&g) + 1;
  gpsm_ctx c;
  memset(&c, 0, sizeof(c));
  c.g = &g;
  c.nv = qq.nv;
  c.out = out;
  c.cap = cap;
  for (int u = 0; u < qq.nv; ++u) {
    for (int j = 0; j < qq.con[u].cnt; ++j)
      if (qq.con[u].v[j] < u) c.lv[u][c.nlv[u]++] = qq.con[u].v[j];
    for (int v = 0; v < u; ++v) cs_push(&c.cond[u], cond_type(qq.order, u, v), v);
    c.buf[u] = (uint32_t*)malloc(md * sizeof(uint32_t));
  }
  c.tmp = (uint32_t*)malloc(md * sizeof(uint32_t));
  uint32_t path[OR_MAXQ];
  for (uint64_t v = v_begin; v < v_end; ++v) {
    path[0] = (uint32_t)v;
    c.level[0]++;
    if (c.nv == 1 && c.out && c.rows < c.cap) c.out[c.rows++] = (uint32_t)v;
    gpsm_dfs(&c, 1, path);
  }
  for (int u = 0; u < qq.nv; ++u) free(c.buf[u]);
  free(c.tmp);
  *rows = c.rows;
  *total = c.level[qq.nv - 1];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* LIGHT plan evaluation.  The op semantics follow
 *   COMPUTE             op/LIGHTCompute.cuh:12-71 + GpsmBinSearch
 *                       (GPSMBinSearch.cuh:146-197): intersect the adjacency
 *                       lists of bconn[u], keep candidates passing ccond[u].
 *   COMPUTE_PATH_COUNT  op/LIGHTComputePathCount.cuh:11-68: same, size only.
 *   MATERIALIZE         op/LIGHTMaterialize.cuh:12-162: level 0 = the start
 *                       vertex; else every candidate passing mcond[u].
 *   FILTER_COMPUTE      op/LIGHTFilterCompute.cuh:22-155: keep candidates
 *                       passing fcond[u].
 *   COMPUTE_COUNT       op/LIGHTComputeCount.cuh:13-119 with
 *                       ThreadCountWithComputedVertex
 *                       (CheckConstraintsCommon.cuh:131-155).
 *   COUNT               op/LIGHTCount.cuh:22-195 (one or two unmaterialized
 *                       vertices) or op/LIGHTDirectCount.cuh:13-122 when the
 *                       compress level is SPECIAL and CanDirectCount holds.
 * A depth-first walk over the plan replaces the reference's breadth-first
 * batches; every op still runs for every partial instance that reaches it. */
typedef struct {
  const graph_t* g;
  const or_plan* p;
  uint32_t M[OR_MAXQ];
  uint32_t* cand[OR_MAXQ];      /* current candidate list per vertex */
  size_t ncand[OR_MAXQ];
  uint32_t* store[OR_MAXOPS];   /* per exec level storage */
  uint32_t* tmp;
  uint64_t level[OR_MAXOPS];
  uint64_t count;
} light_ctx;

/* lower bound, GPUBinSearch.cuh:18-31 */
static size_t lbound(const uint32_t* a, size_t n, uint32_t k) {
  size_t l = 0, r = n;
  while (l < r) { size_t m = (l + r) / 2; if (a[m] < k) l = m + 1; else r = m; }
  return l;
}

/* ThreadCountWithComputedVertex, CheckConstraintsCommon.cuh:131-155:
 * number of elements y of arr with  elem (op) y. */
static uint64_t count_with_computed(const uint32_t* arr, size_t n, uint32_t elem, int op) {
  if (n == 0) return 0;
  size_t pos = lbound(arr, n, elem);
  int eq = pos < n && arr[pos] == elem;
  if (op == OR_LT) return n - (pos + (eq ? 1 : 0));
  if (op == OR_GT) return pos;
  return n - (eq ? 1 : 0);
}

/* GeneralCountTwoComputedVertices, LIGHTCount.cuh:22-116 (per path). */
static uint64_t count_two(const uint32_t* c0, size_t n0, const uint32_t* c1, size_t n1, int op) {
  uint64_t tot = 0;
  int small0 = n0 <= n1;
  const uint32_t* s = small0 ? c0 : c1;
  size_t ns = small0 ? n0 : n1;
  const uint32_t* a = small0 ? c1 : c0;
  size_t na = small0 ? n1 : n0;
  for (size_t r = 0; r < ns; ++r) {
    uint32_t e = s[r];
    size_t pos = lbound(a, na, e);
    int eq = pos < na && a[pos] == e;
    if (op == OR_LT) tot += small0 ? na - (pos + (eq ? 1 : 0)) : pos;
    else if (op == OR_GT) tot += small0 ? pos : na - (pos + (eq ? 1 : 0));
    else tot += na - (eq ? 1 : 0);
  }
  return tot;
}

/* CanDirectCount, LIGHTDirectCount.cuh:13-19 */
static int can_direct(int q) { return q >= 0 && q <= 11 && q != 4 && q != 9; }

/* ThreadDirectCount, LIGHTDirectCount.cuh:21-78 */
static uint64_t direct_count(const uint64_t* b, int q) {
  uint64_t r = 0;
  switch (q) {
    case 0: r = b[2]; break;
    case 1: r = b[3]; break;
    case 2: if (b[1] > 0) r = b[1] * (b[1] - 1) / 2; break;
    case 3: r = b[3]; break;
    case 5:
      r = b[1] * b[3] * b[5];
      if (r > 0) r = r - b[3] * (b[1] - 1 + b[5] - 1) - b[3] * b[3];
      break;
    case 6: if (b[4] > 0 && b[2] > 0) r = b[4] * b[2] - b[2]; break;
    case 7: r = b[4]; break;
    case 8: r = b[4]; break;
    case 10: if (b[2] > 0) r = b[2] * (b[2] - 1) / 2; break;
    case 11: r = b[5]; break;
    default: r = 0; break;
  }
  return r;
}

static int all_ok(const or_cset* cd, uint32_t x, const uint32_t* M) {
  for (int j = 0; j < cd->cnt; ++j)
    if (!cond_ok(cd->op[j], x, M[cd->v[j]])) return 0;
  return 1;
}

/* Intersection of bconn[u] with ccond[u] applied, into out. */
static size_t light_compute(light_ctx* c, int u, uint32_t* out) {
  const or_plan* p = c->p;
  int lv[OR_MAXQ];
  for (int j = 0; j < p->bconn[u].cnt; ++j) lv[j] = p->bconn[u].v[j];
  size_t k = mway(c->g, c->M, lv, p->bconn[u].cnt, out, c->tmp);
  size_t w = 0;
  for (size_t i = 0; i < k; ++i)
    if (all_ok(&p->ccond[u], out[i], c->M)) out[w++] = out[i];
  return w;
}

static void light_run(light_ctx* c, int L) {
  const or_plan* p = c->p;
  if (L == p->nops) return;
  int op = p->op[L], u = p->opv[L];
  switch (op) {
    case OR_COMPUTE: {
      size_t k = light_compute(c, u, c->store[L]);
      uint32_t* sv = c->cand[u]; size_t sn = c->ncand[u];
      c->cand[u] = c->store[L]; c->ncand[u] = k;
      c->level[L] += k;
      light_run(c, L + 1);
      c->cand[u] = sv; c->ncand[u] = sn;
    } break;
    case OR_COMPUTE_PATH_COUNT: {
      size_t k = light_compute(c, u, c->store[L]);
      size_t sn = c->ncand[u];
      c->ncand[u] = k; /* only the size is kept */
      c->level[L] += k;
      light_run(c, L + 1);
      c->ncand[u] = sn;
    } break;
    case OR_MATERIALIZE: {
      /* level > 0; level 0 is handled by the caller */
      const uint32_t* cs = c->cand[u];
      size_t n = c->ncand[u];
      for (size_t i = 0; i < n; ++i) {
        uint32_t x = cs[i];
        if (!all_ok(&p->mcond[u], x, c->M)) continue;
        c->M[u] = x;
        c->level[L]++;
        light_run(c, L + 1);
      }
    } break;
    case OR_FILTER_COMPUTE: {
      const uint32_t* cs = c->cand[u];
      size_t n = c->ncand[u], w = 0;
      uint32_t* dst = c->store[L];
      for (size_t i = 0; i < n; ++i)
        if (all_ok(&p->fcond[u], cs[i], c->M)) dst[w++] = cs[i];
      uint32_t* sv = c->cand[u]; size_t sn = c->ncand[u];
      c->cand[u] = dst; c->ncand[u] = w;
      c->level[L] += w;
      light_run(c, L + 1);
      c->cand[u] = sv; c->ncand[u] = sn;
    } break;
    case OR_COMPUTE_COUNT: {
      size_t k = light_compute(c, u, c->store[L]);
      uint64_t add = 0;
      const or_vset* un = &p->cun[L];
      if (un->cnt == 0) {
        add = k;
      } else {
        int w = un->v[0];
        int opw = cond_type(p->order, u, w);
        for (size_t i = 0; i < k; ++i)
          add += count_with_computed(c->cand[w], c->ncand[w], c->store[L][i], opw);
      }
      c->level[L] += add;
      c->count += add;
    } break;
    case OR_COUNT: {
      const or_vset* un = &p->cun[L];
      uint64_t add = 0;
      if (p->compress_level == OR_CL_SPECIAL && can_direct(p->q)) {
        uint64_t b[OR_MAXQ] = {0};
        for (int j = 0; j < un->cnt; ++j) b[un->v[j]] = c->ncand[un->v[j]];
        add = direct_count(b, p->q);
      } else if (un->cnt == 1) {
        add = c->ncand[un->v[0]];
      } else if (un->cnt == 2) {
        int u0 = un->v[0], u1 = un->v[1];
        add = count_two(c->cand[u0], c->ncand[u0], c->cand[u1], c->ncand[u1],
                        cond_type(p->order, u0, u1));
      }
      c->level[L] += add;
      c->count += add;
    } break;
  }
}

static int light_count_impl(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                            int compress_level, int enable_ordering, uint32_t v_begin,
                            uint32_t v_end, int nthreads, uint64_t* out_count, uint64_t* level_out,
                            uint64_t* per_vertex);

int or_light_count(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                   int compress_level, int enable_ordering, uint32_t v_begin, uint32_t v_end,
                   int nthreads, uint64_t* out_count, uint64_t* level_out) {
  return light_count_impl(n, row_ptrs, cols, q, compress_level, enable_ordering, v_begin, v_end,
                          nthreads, out_count, level_out, NULL);
}

int or_light_count_per_vertex(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                              int compress_level, int enable_ordering, uint32_t v_begin,
                              uint32_t v_end, int nthreads, uint64_t* per_vertex) {
  uint64_t total;
  if (v_end > n) return -2;
  return light_count_impl(n, row_ptrs, cols, q, compress_level, enable_ordering, v_begin, v_end,
                          nthreads, &total, NULL, per_vertex);
}

static int light_count_impl(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                            int compress_level, int enable_ordering, uint32_t v_begin,
                            uint32_t v_end, int nthreads, uint64_t* out_count, uint64_t* level_out,
                            uint64_t* per_vertex) {
  or_plan p;
  if (or_light_plan(q, compress_level, enable_ordering, &p)) return -1;
  if (v_end > n) v_end = (uint32_t)n;
  graph_t g = {n, row_ptrs, cols};
  uint64_t md = max_degree(&g) + 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  uint64_t lv_tot[OR_MAXOPS] = {0};
  uint64_t total = 0;
#pragma omp parallel num_threads(nthreads)
  {
    light_ctx c;
    memset(&c, 0, sizeof(c));
    c.g = &g;
    c.p = &p;
    for (int L = 0; L < p.nops; ++L) c.store[L] = (uint32_t*)malloc(md * sizeof(uint32_t));
    c.tmp = (uint32_t*)malloc(md * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 64)
    for (int64_t v = v_begin; v < (int64_t)v_end; ++v) {
      const uint64_t before = c.count;
      c.M[p.seq[0]] = (uint32_t)v; /* InitFirstLevel, LIGHTMaterialize.cuh:12-21 */
      c.level[0]++;
      light_run(&c, 1);
      if (per_vertex) per_vertex[v - v_begin] = c.count - before;
    }
#pragma omp critical
    {
      for (int L = 0; L < p.nops; ++L) lv_tot[L] += c.level[L];
      total += c.count;
    }
    for (int L = 0; L < p.nops; ++L) free(c.store[L]);
    free(c.tmp);
  }
  *out_count = total;
  if (level_out) for (int L = 0; L < p.nops; ++L) level_out[L] = lv_tot[L];
  return 0;
}
