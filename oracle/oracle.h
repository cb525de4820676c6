/*
 * oracle.h -- CPU restatement of the reference's subgraph-enumeration path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product (subgraphmatchgpu_amd, libsme.so) never links, loads or calls
 * anything here.
 *
 * What it restates (file:line are in the reference tree, src/...):
 *   - the query registry             query/common/Query.h:84-433, Meta.h:104-126
 *   - the LIGHT plan generator       query/optimizer/LazyTraversalPlanGenerator.h:11-300,
 *                                    query/optimizer/LazyTraversalPlanUtils.h:40-158,
 *                                    query/optimizer/LazyTraversalPlanHardCode.h:59-117,
 *                                    query/LazyTraversalPlan.h:13-29,173-197,
 *                                    query/common/Plan.h:32-43,60-88,202-217
 *   - the CPU WCOJ DFS count         cpu/CPUWCOJ.h:20-111, cpu/common/CPUFilter.h:7-34,
 *                                    cpu/common/CPUIntersection.h:13-33,64-99
 *   - the LIGHT op semantics, as a   gpu/impl/light/op (see oracle.c per op)
 *     depth-first evaluation of the
 *     plan with per-op output counts
 *
 * Pinning status is recorded at the top of oracle.c.
 */
#ifndef SME_ORACLE_H
#define SME_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXQ 6   /* kMaxQueryVerticesNum, Meta.h:30 */
#define OR_MAXOPS 16

/* LazyTraversalOperation, query/common/LazyTraversalCommon.h:11-31 */
enum { OR_COMPUTE = 0, OR_COMPUTE_PATH_COUNT, OR_MATERIALIZE, OR_FILTER_COMPUTE,
       OR_COUNT, OR_COMPUTE_COUNT };
/* CondOperator, Meta.h:126 */
enum { OR_LT = 0, OR_GT, OR_NE, OR_NONE };
/* LazyTraversalCompressLevel, LazyTraversalCommon.h:38-76 */
enum { OR_CL_MATERIALIZE = 0, OR_CL_NON_MATERIALIZE, OR_CL_NON_MATERIALIZE_OPT,
       OR_CL_SPECIAL };

typedef struct { int cnt; int v[OR_MAXQ]; } or_vset;
typedef struct { int cnt; int op[OR_MAXQ]; int v[OR_MAXQ]; } or_cset;

typedef struct {
  int q;
  int nv;
  or_vset con[OR_MAXQ];
  or_cset order[OR_MAXQ];
} or_query;

typedef struct {
  int q, compress_level, enable_ordering;
  int nv;
  int seq[OR_MAXQ];
  int nops;
  int op[OR_MAXOPS];
  int opv[OR_MAXOPS];
  or_vset mat[OR_MAXOPS];     /* materialized before exec level i */
  or_vset cun[OR_MAXOPS];     /* computed but unmaterialized before level i */
  or_vset bconn[OR_MAXQ];     /* backward connectivity by match order */
  or_cset ccond[OR_MAXQ];     /* compute condition */
  or_cset mcond[OR_MAXQ];     /* materialize condition */
  or_cset fcond[OR_MAXQ];     /* filter condition */
  or_cset order[OR_MAXQ];     /* the query's (unrelabelled) ordering */
} or_plan;

int or_query_init(int q, int enable_ordering, or_query* out);
int or_light_plan(int q, int compress_level, int enable_ordering, or_plan* out);
/* Canonical text form of a plan, shared with the reference-side dumper. */
int or_plan_text(const or_plan* p, char* buf, size_t cap);

/* CPU_WCOJ count over start vertices [v_begin, v_end) of pattern vertex 0. */
int or_wcoj_count(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols,
                  int q, int enable_ordering, uint32_t v_begin, uint32_t v_end,
                  int nthreads, uint64_t* out_count);

/* CPU_WCOJ count over an explicit list of start vertices. */
int or_wcoj_count_list(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                       int enable_ordering, const uint32_t* starts, uint64_t nstarts,
                       int nthreads, uint64_t* out_count);

/* LIGHT-plan evaluation over start vertices [v_begin, v_end) of seq[0].
 * level_out[i] (i < nops) receives the summed output size of exec level i:
 * paths for MATERIALIZE, candidates for COMPUTE / FILTER_COMPUTE /
 * COMPUTE_PATH_COUNT, the count contribution for COUNT / COMPUTE_COUNT. */
int or_light_count(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols,
                   int q, int compress_level, int enable_ordering,
                   uint32_t v_begin, uint32_t v_end, int nthreads,
                   uint64_t* out_count, uint64_t* level_out);

/* or_light_count split by start vertex: per_vertex[v - v_begin] receives the
 * count of the embeddings whose first matched vertex is v (v_end <= n). */
int or_light_count_per_vertex(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                              int compress_level, int enable_ordering, uint32_t v_begin,
                              uint32_t v_end, int nthreads, uint64_t* per_vertex);

/* GPSM semantics (gpu/impl/gpsm/HybridGPSMGPUProcessor.cuh:23-205): per
 * level, intersect the ordered connectivity with whole-ordered conditions and
 * materialize; level_out[l] = partial instances after level l (l < nv). */
int or_gpsm_count(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols,
                  int q, int enable_ordering, uint32_t v_begin, uint32_t v_end,
                  int nthreads, uint64_t* out_count, uint64_t* level_out);
/* GPSM match list, lexicographic; up to cap rows of nv u32 into out. */
int or_gpsm_list(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int q,
                 int enable_ordering, uint32_t v_begin, uint32_t v_end, uint32_t* out,
                 uint64_t cap, uint64_t* rows, uint64_t* total);

/* Whole-graph pattern counts by degree-ordered clique listing (census.c):
 * an independent check of the LIGHT counts on large graphs, sharing no code
 * or ordering with the functions above.  kmax (3..5) is the largest clique
 * listed; triangles and diamonds are always counted. */
typedef struct {
  uint64_t edges;           /* undirected edges (self loops excluded) */
  uint64_t self_loops;
  uint64_t triangles;       /* Q0 */
  uint64_t diamonds;        /* Q2: sum over edges of C(t(e), 2) */
  uint64_t k4;              /* Q3 (kmax >= 4) */
  uint64_t k5;              /* Q7 (kmax >= 5) */
  uint64_t max_out_degree;  /* of the degree orientation */
} or_census_t;
int or_census(uint64_t n, const uint64_t* row_ptrs, const uint32_t* cols, int kmax, int nthreads,
              or_census_t* out);

#ifdef __cplusplus
}
#endif
#endif
