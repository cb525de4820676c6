/*
 * census.c -- pattern counts of a whole graph by textbook algorithms that
 * share nothing with the LIGHT plan.
 *
 * TEST INFRASTRUCTURE ONLY (like the rest of oracle/): tests/ and the fixture
 * scripts under tests/golden use it to pin counts on graphs too large for the
 * plan-following restatement in oracle.c.  The product never loads it.
 *
 * Method (degree-ordered listing, Chiba-Nishizeki / Latapy):
 *   - vertices are ranked by (degree, id); every edge is oriented from the
 *     lower to the higher rank, so each clique is listed exactly once, from
 *     its lowest-ranked vertex, by sorted merges of out-lists;
 *   - triangles: sum over oriented edges u->v of |N+(u) & N+(v)|;
 *   - per-edge support t(e) = triangles through e (each listed triangle adds
 *     one to each of its three edges);
 *   - diamonds (4-cycle with one chord): every diamond is a chord edge plus
 *     an unordered pair of its common neighbours, so the count is
 *     sum_e t(e)(t(e)-1)/2;
 *   - 4-cliques and 5-cliques: one and two more merge levels.
 * The LIGHT plan instead orders by vertex id, picks pivots per path and
 * counts the diamond with its direct closed form per (v0, v2) pair; the two
 * agree only if both are right.  Self loops are ignored (they never take
 * part in a simple pattern).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* Merge of sorted a[0,na) and b[0,nb); out receives the common values and,
 * when ia / ib are given, their positions in a and b. */
static size_t isect_pos(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out,
                        uint32_t* ia, uint32_t* ib) {
  size_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) {
      ++i;
    } else if (a[i] > b[j]) {
      ++j;
    } else {
      if (out) out[k] = a[i];
      if (ia) ia[k] = (uint32_t)i;
      if (ib) ib[k] = (uint32_t)j;
      ++k;
      ++i;
      ++j;
    }
  }
  return k;
}

static size_t isect_count(const uint32_t* a, size_t na, const uint32_t* b, size_t nb) {
  size_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    const uint32_t x = a[i], y = b[j];
    i += x <= y;
    j += y <= x;
    k += x == y;
  }
  return k;
}

int or_census(uint64_t n, const uint64_t* rp, const uint32_t* cols, int kmax, int nthreads,
              or_census_t* out) {
  if (!rp || !out || kmax < 3 || kmax > 5) return 1;
  memset(out, 0, sizeof(*out));
  if (n == 0) return 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  /* out-lists: neighbours of higher (degree, id) rank, in id order */
  uint64_t* orp = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
  if (!orp) return 2;
  orp[0] = 0;
  uint64_t loops = 0;
  for (uint64_t u = 0; u < n; ++u) {
    const uint64_t du = rp[u + 1] - rp[u];
    uint64_t k = 0;
    for (uint64_t j = rp[u]; j < rp[u + 1]; ++j) {
      const uint32_t v = cols[j];
      if (v == u) {
        ++loops;
        continue;
      }
      const uint64_t dv = rp[v + 1] - rp[v];
      if (dv > du || (dv == du && v > u)) ++k;
    }
    orp[u + 1] = orp[u] + k;
  }
  const uint64_t m = orp[n];
  uint32_t* oc = (uint32_t*)malloc((m ? m : 1) * sizeof(uint32_t));
  uint32_t* sup = (uint32_t*)calloc(m ? m : 1, sizeof(uint32_t));
  if (!oc || !sup) {
    free(orp);
    free(oc);
    free(sup);
    return 2;
  }
  uint64_t maxout = 0;
  for (uint64_t u = 0; u < n; ++u) {
    const uint64_t du = rp[u + 1] - rp[u];
    uint64_t k = orp[u];
    for (uint64_t j = rp[u]; j < rp[u + 1]; ++j) {
      const uint32_t v = cols[j];
      if (v == u) continue;
      const uint64_t dv = rp[v + 1] - rp[v];
      if (dv > du || (dv == du && v > u)) oc[k++] = v;
    }
    if (orp[u + 1] - orp[u] > maxout) maxout = orp[u + 1] - orp[u];
  }
  uint64_t tri = 0, k4 = 0, k5 = 0;
  int fail = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : tri, k4, k5)
  {
    const size_t cap = maxout ? maxout : 1;
    uint32_t* b1 = (uint32_t*)malloc(cap * sizeof(uint32_t));
    uint32_t* b2 = (uint32_t*)malloc(cap * sizeof(uint32_t));
    uint32_t* ia = (uint32_t*)malloc(cap * sizeof(uint32_t));
    uint32_t* ib = (uint32_t*)malloc(cap * sizeof(uint32_t));
    if (!b1 || !b2 || !ia || !ib) {
#pragma omp atomic write
      fail = 1;
    } else {
#pragma omp for schedule(dynamic, 16)
      for (uint64_t u = 0; u < n; ++u) {
        const uint32_t* A = oc + orp[u];
        const size_t na = orp[u + 1] - orp[u];
        for (size_t iv = 0; iv < na; ++iv) {
          const uint32_t v = A[iv];
          const uint32_t* Bv = oc + orp[v];
          const size_t nbv = orp[v + 1] - orp[v];
          const size_t nb = isect_pos(A, na, Bv, nbv, b1, ia, ib);
          tri += nb;
          if (nb) {
#pragma omp atomic
            sup[orp[u] + iv] += (uint32_t)nb;
          }
          for (size_t k = 0; k < nb; ++k) {
#pragma omp atomic
            sup[orp[u] + ia[k]] += 1u;
#pragma omp atomic
            sup[orp[v] + ib[k]] += 1u;
          }
          if (kmax < 4) continue;
          for (size_t iw = 0; iw < nb; ++iw) {
            const uint32_t w = b1[iw];
            const uint32_t* Bw = oc + orp[w];
            const size_t nbw = orp[w + 1] - orp[w];
            if (kmax < 5) {
              k4 += isect_count(b1, nb, Bw, nbw);
              continue;
            }
            const size_t nc = isect_pos(b1, nb, Bw, nbw, b2, NULL, NULL);
            k4 += nc;
            for (size_t ix = 0; ix < nc; ++ix) {
              const uint32_t x = b2[ix];
              k5 += isect_count(b2, nc, oc + orp[x], orp[x + 1] - orp[x]);
            }
          }
        }
      }
    }
    free(b1);
    free(b2);
    free(ia);
    free(ib);
  }
  if (fail) {
    free(orp);
    free(oc);
    free(sup);
    return 2;
  }
  uint64_t dia = 0;
  for (uint64_t e = 0; e < m; ++e) dia += (uint64_t)sup[e] * (sup[e] ? sup[e] - 1 : 0) / 2;
  out->edges = m;
  out->self_loops = loops;
  out->triangles = tri;
  out->diamonds = dia;
  out->k4 = kmax >= 4 ? k4 : 0;
  out->k5 = kmax >= 5 ? k5 : 0;
  out->max_out_degree = maxout;
  free(orp);
  free(oc);
  free(sup);
  return 0;
}
