// graph.cpp -- see graph.h.
#include "graph.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <sstream>

namespace sme {

uint64_t SplitMix64::Next() {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

Xoshiro256ss::Xoshiro256ss(uint64_t seed) {
  SplitMix64 sm{seed};
  for (auto& w : s) w = sm.Next();
}

static inline uint64_t Rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Xoshiro256ss::Next() {
  const uint64_t result = Rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = Rotl(s[3], 45);
  return result;
}

double Xoshiro256ss::Uniform() { return static_cast<double>(Next() >> 11) * 0x1.0p-53; }

Status ReadBin(const std::string& path, HostGraph* g, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    *err = "cannot open " + path;
    return kErrIo;
  }
  uint64_t hdr[4];
  if (std::fread(hdr, sizeof(uint64_t), 4, f) != 4) {
    std::fclose(f);
    *err = "short BIN header in " + path;
    return kErrIo;
  }
  if (hdr[0] != sizeof(uintV) || hdr[1] != sizeof(uintE)) {
    std::fclose(f);
    *err = "BIN id widths are not (4, 8) in " + path;
    return kErrInvalid;
  }
  // the header's sizes must match the file before anything is allocated
  if (std::fseek(f, 0, SEEK_END) != 0) {
    std::fclose(f);
    *err = "cannot seek in " + path;
    return kErrIo;
  }
  const long fsize = std::ftell(f);
  if (fsize < 0 || std::fseek(f, 4 * sizeof(uint64_t), SEEK_SET) != 0) {
    std::fclose(f);
    *err = "cannot seek in " + path;
    return kErrIo;
  }
  const uint64_t body = static_cast<uint64_t>(fsize) - 4 * sizeof(uint64_t);
  if (hdr[2] >= 0xffffffffull || hdr[3] > body / sizeof(uintV) ||
      (hdr[2] + 1) * sizeof(uintE) + hdr[3] * sizeof(uintV) != body) {
    std::fclose(f);
    *err = "BIN header sizes do not match the file size of " + path;
    return kErrInvalid;
  }
  g->n = hdr[2];
  g->nnz = hdr[3];
  g->row_ptrs.resize(g->n + 1);
  g->cols.resize(g->nnz);
  size_t got = std::fread(g->row_ptrs.data(), sizeof(uintE), g->n + 1, f);
  got += std::fread(g->cols.data(), sizeof(uintV), g->nnz, f);
  const bool trailing = std::fgetc(f) != EOF;
  std::fclose(f);
  if (got != g->n + 1 + g->nnz || trailing) {
    *err = "BIN body size mismatch in " + path;
    return kErrIo;
  }
  return kOk;
}

Status WriteBin(const std::string& path, const HostGraph& g, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    *err = "cannot create " + path;
    return kErrIo;
  }
  const uint64_t hdr[4] = {sizeof(uintV), sizeof(uintE), g.n, g.nnz};
  size_t put = std::fwrite(hdr, sizeof(uint64_t), 4, f);
  put += std::fwrite(g.row_ptrs.data(), sizeof(uintE), g.n + 1, f);
  put += std::fwrite(g.cols.data(), sizeof(uintV), g.nnz, f);
  const bool ok = std::fclose(f) == 0 && put == 4 + g.n + 1 + g.nnz;
  if (!ok) {
    *err = "write failed for " + path;
    return kErrIo;
  }
  return kOk;
}

Status ReadSnap(const std::string& path, bool directed, HostGraph* g, std::string* err) {
  std::ifstream in(path);
  if (!in) {
    *err = "cannot open " + path;
    return kErrIo;
  }
  std::vector<std::pair<uint64_t, uint64_t>> edges;
  uint64_t lo = UINT64_MAX, hi = 0;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || !std::isdigit(static_cast<unsigned char>(line[0]))) continue;
    std::istringstream is(line);
    uint64_t a, b;
    if (!(is >> a >> b)) continue;
    lo = std::min(lo, std::min(a, b));
    hi = std::max(hi, std::max(a, b));
    edges.emplace_back(a, b);
  }
  if (edges.empty()) {
    g->n = 0;
    g->nnz = 0;
    g->row_ptrs.assign(1, 0);
    g->cols.clear();
    return kOk;
  }
  g->n = hi - lo + 1;
  std::vector<uint64_t> deg(g->n + 1, 0);
  for (auto& e : edges) {
    e.first -= lo;
    e.second -= lo;
    deg[e.first]++;
    if (!directed) deg[e.second]++;
  }
  g->row_ptrs.assign(g->n + 1, 0);
  for (uint64_t u = 0; u < g->n; ++u) g->row_ptrs[u + 1] = g->row_ptrs[u] + deg[u];
  g->nnz = g->row_ptrs[g->n];
  g->cols.resize(g->nnz);
  std::vector<uint64_t> fill(g->row_ptrs.begin(), g->row_ptrs.end() - 1);
  for (auto& e : edges) {
    g->cols[fill[e.first]++] = static_cast<uintV>(e.second);
    if (!directed) g->cols[fill[e.second]++] = static_cast<uintV>(e.first);
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < static_cast<int64_t>(g->n); ++u)
    std::sort(g->cols.begin() + g->row_ptrs[u], g->cols.begin() + g->row_ptrs[u + 1]);
  return kOk;
}

bool CheckCsr(uint64_t n, uint64_t nnz, const uintE* rp, const uintV* cols, bool* self_loops,
              std::string* err) {
  if (rp[0] != 0 || rp[n] != nnz) {
    *err = "row_ptrs[0] must be 0 and row_ptrs[n] must equal nnz";
    return false;
  }
  bool bad_order = false, bad_id = false, loops = false;
#pragma omp parallel for schedule(dynamic, 4096) reduction(|| : bad_order, bad_id, loops)
  for (int64_t u = 0; u < static_cast<int64_t>(n); ++u) {
    if (rp[u] > rp[u + 1]) {
      bad_order = true;
      continue;
    }
    for (uint64_t j = rp[u]; j < rp[u + 1]; ++j) {
      if (cols[j] >= n) bad_id = true;
      if (j > rp[u] && cols[j] <= cols[j - 1]) bad_order = true;
      if (cols[j] == static_cast<uintV>(u)) loops = true;
    }
  }
  if (bad_id) {
    *err = "a column id is >= n";
    return false;
  }
  if (bad_order) {
    *err = "adjacency lists must be strictly increasing (sorted, no parallel edges)";
    return false;
  }
  *self_loops = loops;
  return true;
}

void BuildSimpleUndirected(uint64_t n, const std::vector<uint64_t>& packed, HostGraph* g) {
  // counting pass: degree with both directions, self loops dropped
  const int nt = omp_get_max_threads();
  std::vector<uint64_t> deg(n + 1, 0);
  {
    std::vector<std::vector<uint64_t>> local(nt);
#pragma omp parallel num_threads(nt)
    {
      auto& d = local[omp_get_thread_num()];
      d.assign(n, 0);
#pragma omp for schedule(static)
      for (int64_t i = 0; i < static_cast<int64_t>(packed.size()); ++i) {
        const uint64_t a = packed[i] >> 32, b = packed[i] & 0xffffffffull;
        if (a == b) continue;
        d[a]++;
        d[b]++;
      }
    }
    for (auto& d : local)
      for (uint64_t u = 0; u < n; ++u) deg[u] += d[u];
  }
  std::vector<uint64_t> off(n + 1, 0);
  for (uint64_t u = 0; u < n; ++u) off[u + 1] = off[u] + deg[u];
  std::vector<uintV> raw(off[n]);
  std::vector<uint64_t> cur(off.begin(), off.end() - 1);
  for (uint64_t e : packed) {
    const uint64_t a = e >> 32, b = e & 0xffffffffull;
    if (a == b) continue;
    raw[cur[a]++] = static_cast<uintV>(b);
    raw[cur[b]++] = static_cast<uintV>(a);
  }
  // sort + unique per vertex, then compact
  std::vector<uint64_t> kept(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < static_cast<int64_t>(n); ++u) {
    auto b = raw.begin() + off[u], e = raw.begin() + off[u + 1];
    std::sort(b, e);
    kept[u] = static_cast<uint64_t>(std::unique(b, e) - b);
  }
  g->n = n;
  g->row_ptrs.assign(n + 1, 0);
  for (uint64_t u = 0; u < n; ++u) g->row_ptrs[u + 1] = g->row_ptrs[u] + kept[u];
  g->nnz = g->row_ptrs[n];
  g->cols.resize(g->nnz);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < static_cast<int64_t>(n); ++u)
    std::copy(raw.begin() + off[u], raw.begin() + off[u] + kept[u], g->cols.begin() + g->row_ptrs[u]);
}

void GenerateRmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                  HostGraph* g) {
  const uint64_t n = 1ull << scale;
  const uint64_t m = static_cast<uint64_t>(edge_factor) << scale;
  constexpr uint64_t kChunk = 1ull << 16;  // draws per independent stream
  const uint64_t nchunks = (m + kChunk - 1) / kChunk;
  std::vector<uint64_t> packed(m);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t ch = 0; ch < static_cast<int64_t>(nchunks); ++ch) {
    SplitMix64 mix{seed ^ (0xD1B54A32D192ED03ull * static_cast<uint64_t>(ch + 1))};
    Xoshiro256ss rng(mix.Next());
    const uint64_t lo = ch * kChunk, hi = std::min(m, lo + kChunk);
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t src = 0, dst = 0;
      for (int bit = 0; bit < scale; ++bit) {
        const double r = rng.Uniform();
        const uint64_t row = r >= ab ? 1 : 0;                      // quadrants c, d
        const uint64_t col = (r >= a && r < ab) || r >= abc ? 1 : 0;  // quadrants b, d
        src = (src << 1) | row;
        dst = (dst << 1) | col;
      }
      packed[i] = (src << 32) | dst;
    }
  }
  BuildSimpleUndirected(n, packed, g);
}

namespace {

// First m distinct unordered pairs of the stream (x, y) = (Next() % n,
// Next() % n), x != y, as sorted keys (min << 32) | max.  Drawing exactly the
// current deficit per round and de-duplicating keeps the set identical to a
// one-at-a-time rejection loop over the same stream.
void DrawDistinctPairs(uint64_t n, uint64_t m, Xoshiro256ss* rng, std::vector<uint64_t>* keys) {
  keys->clear();
  std::vector<uint64_t> fresh, merged;
  while (keys->size() < m) {
    const uint64_t want = m - keys->size();
    fresh.clear();
    for (uint64_t i = 0; i < want; ++i) {
      uint64_t x = rng->Next() % n;
      uint64_t y = rng->Next() % n;
      if (x == y) continue;
      if (x > y) std::swap(x, y);
      fresh.push_back((x << 32) | y);
    }
    std::sort(fresh.begin(), fresh.end());
    fresh.erase(std::unique(fresh.begin(), fresh.end()), fresh.end());
    merged.clear();
    std::set_union(keys->begin(), keys->end(), fresh.begin(), fresh.end(),
                   std::back_inserter(merged));
    keys->swap(merged);
  }
}

// CSR of the simple graph given by sorted distinct keys (x << 32) | y, x < y,
// written at vertex offset `vbase` into rp[0..n) (entry offsets starting at
// `ebase`; the caller sets the end offset) and cols.  Filling rows in key order leaves every list sorted:
// neighbours below u are placed before those above it, each in order.
void FillSortedPairs(uint64_t n, const std::vector<uint64_t>& keys, uint64_t vbase, uint64_t ebase,
                     uintE* rp, uintV* cols) {
  std::vector<uint64_t> cur(n + 1, 0);
  for (uint64_t k : keys) {
    cur[(k >> 32) + 1]++;
    cur[(k & 0xffffffffull) + 1]++;
  }
  for (uint64_t u = 0; u < n; ++u) cur[u + 1] += cur[u];
  for (uint64_t u = 0; u < n; ++u) rp[u] = ebase + cur[u];
  for (uint64_t k : keys) {
    const uint64_t x = k >> 32, y = k & 0xffffffffull;
    cols[ebase + cur[x]++] = static_cast<uintV>(vbase + y);
    cols[ebase + cur[y]++] = static_cast<uintV>(vbase + x);
  }
}

}  // namespace

void GenerateErdosRenyi(uint64_t n, uint64_t m, uint64_t seed, HostGraph* g) {
  const uint64_t max_edges = n * (n - 1) / 2;
  if (m > max_edges) m = max_edges;
  Xoshiro256ss rng(seed);
  std::vector<uint64_t> keys;
  DrawDistinctPairs(n, m, &rng, &keys);
  g->n = n;
  g->nnz = 2 * keys.size();
  g->row_ptrs.assign(n + 1, 0);
  g->cols.resize(g->nnz);
  FillSortedPairs(n, keys, 0, 0, g->row_ptrs.data(), g->cols.data());
  g->row_ptrs[n] = g->nnz;
}

void GenerateErdosRenyiBatch(uint64_t ngraphs, uint64_t n_min, uint64_t n_max, double deg_min,
                             double deg_max, uint64_t seed, HostGraph* g,
                             std::vector<uint64_t>* graph_offs) {
  // parameters of every member from one master stream, in member order
  std::vector<uint64_t> nk(ngraphs), mk(ngraphs), sk(ngraphs);
  Xoshiro256ss master(seed);
  for (uint64_t k = 0; k < ngraphs; ++k) {
    nk[k] = n_min + master.Next() % (n_max - n_min + 1);
    const double deg = deg_min + (deg_max - deg_min) * master.Uniform();
    const uint64_t max_edges = nk[k] * (nk[k] - 1) / 2;
    mk[k] = std::min<uint64_t>(max_edges, static_cast<uint64_t>(std::llround(nk[k] * deg / 2.0)));
    sk[k] = master.Next();
  }
  graph_offs->assign(ngraphs + 1, 0);
  std::vector<uint64_t> eoffs(ngraphs + 1, 0);
  for (uint64_t k = 0; k < ngraphs; ++k) {
    (*graph_offs)[k + 1] = (*graph_offs)[k] + nk[k];
    eoffs[k + 1] = eoffs[k] + 2 * mk[k];
  }
  g->n = (*graph_offs)[ngraphs];
  g->nnz = eoffs[ngraphs];
  g->row_ptrs.assign(g->n + 1, 0);
  g->cols.resize(g->nnz);
#pragma omp parallel
  {
    std::vector<uint64_t> keys;
#pragma omp for schedule(dynamic, 4)
    for (int64_t k = 0; k < static_cast<int64_t>(ngraphs); ++k) {
      Xoshiro256ss rng(sk[k]);
      DrawDistinctPairs(nk[k], mk[k], &rng, &keys);
      FillSortedPairs(nk[k], keys, (*graph_offs)[k], eoffs[k], g->row_ptrs.data() + (*graph_offs)[k],
                      g->cols.data());
    }
  }
  g->row_ptrs[g->n] = g->nnz;
}

void RankByDegree(uint64_t n, const uintE* rp, const uintV* cols, std::vector<uintE>* rrp,
                  std::vector<uintV>* rcols, std::vector<uintV>* new_id) {
  // stable counting sort of the vertices by degree
  uint64_t maxd = 0;
  for (uint64_t v = 0; v < n; ++v) maxd = std::max<uint64_t>(maxd, rp[v + 1] - rp[v]);
  std::vector<uint64_t> start(maxd + 2, 0);
  for (uint64_t v = 0; v < n; ++v) ++start[rp[v + 1] - rp[v] + 1];
  for (uint64_t d = 0; d <= maxd; ++d) start[d + 1] += start[d];
  new_id->resize(n);
  std::vector<uintV> old(n);
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t u = start[rp[v + 1] - rp[v]]++;
    (*new_id)[v] = static_cast<uintV>(u);
    old[u] = static_cast<uintV>(v);
  }
  rrp->assign(n + 1, 0);
  for (uint64_t u = 0; u < n; ++u) (*rrp)[u + 1] = (*rrp)[u] + (rp[old[u] + 1] - rp[old[u]]);
  rcols->resize(rp[n]);
  const uintV* nid = new_id->data();
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < static_cast<int64_t>(n); ++u) {
    const uint64_t v = old[u];
    uintV* out = rcols->data() + (*rrp)[u];
    const uint64_t d = rp[v + 1] - rp[v];
    for (uint64_t i = 0; i < d; ++i) out[i] = nid[cols[rp[v] + i]];
    std::sort(out, out + d);
  }
}

}  // namespace sme
