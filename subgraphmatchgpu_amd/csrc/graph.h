// graph.h -- host CSR graphs: the reference's BIN format, SNAP text input,
// and the seeded synthetic generators used for the benchmark workloads.
//
// BIN layout (src/graph/GraphIO.h:88-136 read, :356-387 write):
//   u64 sizeof(uintV)=4 | u64 sizeof(uintE)=8 | u64 n | u64 nnz |
//   u64 row_ptrs[n+1] | u32 cols[nnz]
// nnz counts directed entries (an undirected edge appears twice).
#pragma once

#include <string>
#include <vector>

#include "common.h"

namespace sme {

struct HostGraph {
  uint64_t n = 0;
  uint64_t nnz = 0;
  std::vector<uintE> row_ptrs;  // n + 1
  std::vector<uintV> cols;      // nnz
};

// Status-returning IO (no asserts, unlike GraphIO.h).
Status ReadBin(const std::string& path, HostGraph* g, std::string* err);
Status WriteBin(const std::string& path, const HostGraph& g, std::string* err);
// SNAP edge list "u v" per line (GraphIO.h:232-311): ids shifted by the
// minimum id, both directions when undirected, lists sorted.
Status ReadSnap(const std::string& path, bool directed, HostGraph* g, std::string* err);

// Structural checks used by the engine: row_ptrs monotone with
// row_ptrs[n] == nnz, every id < n, every list strictly increasing.
// `self_loops` reports whether any u appears in its own list.
bool CheckCsr(uint64_t n, uint64_t nnz, const uintE* row_ptrs, const uintV* cols,
              bool* self_loops, std::string* err);

// Undirected simple graph from an edge list: self loops dropped, both
// directions added, duplicates removed, lists sorted.  Ids stay 0..n-1 and
// isolated vertices are kept.
void BuildSimpleUndirected(uint64_t n, const std::vector<uint64_t>& packed_edges, HostGraph* g);

// Seeded R-MAT: edge_factor * 2^scale draws, each of `scale` quadrant
// choices with probabilities (a, b, c, 1-a-b-c); random stream SplitMix64 ->
// xoshiro256**, one stream per fixed chunk of draws, so the result does not
// depend on the thread count.
void GenerateRmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                  HostGraph* g);

// Seeded Erdos-Renyi G(n, m): m distinct undirected edges drawn by rejection.
void GenerateErdosRenyi(uint64_t n, uint64_t m, uint64_t seed, HostGraph* g);

// `ngraphs` independent G(n, m) graphs packed back to back into one CSR (a
// disjoint union).  Member k has n_k uniform in [n_min, n_max] vertices and
// m_k = round(n_k * d_k / 2) edges for an average degree d_k uniform in
// [deg_min, deg_max); (n_k, d_k, seed_k) come from one master xoshiro256**
// stream in member order, and member k is GenerateErdosRenyi(n_k, m_k,
// seed_k) shifted by graph_offs[k].  graph_offs gets ngraphs + 1 vertex
// offsets.
void GenerateErdosRenyiBatch(uint64_t ngraphs, uint64_t n_min, uint64_t n_max, double deg_min,
                             double deg_max, uint64_t seed, HostGraph* g,
                             std::vector<uint64_t>* graph_offs);

// Degree-ranked copy of a CSR: vertex v becomes new_id[v], its position in
// increasing (degree, id) order, and every list is mapped and re-sorted.  The
// copy is isomorphic to the input, so a pattern counted over all start
// vertices has the same count; listing a clique from its lowest id then
// starts at its lowest-degree vertex, and the lists "above" a vertex hold
// only neighbours of at least its degree (the orientation of
// Chiba-Nishizeki listing).
void RankByDegree(uint64_t n, const uintE* row_ptrs, const uintV* cols, std::vector<uintE>* rrp,
                  std::vector<uintV>* rcols, std::vector<uintV>* new_id);

// Deterministic 64-bit PRNG pair used by the generators.
struct SplitMix64 {
  uint64_t s;
  uint64_t Next();
};
struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed);
  uint64_t Next();
  double Uniform();  // [0, 1) with 53 random bits
};

}  // namespace sme
