// Test driver for the CPU build of the engine: runs one plan through the C
// ABI and prints the count and the per-level outputs.
//   emu_driver GRAPH.bin ALGO QUERY COMPRESS ORDERING V_BEGIN V_END BUDGET_BYTES
// Output: "count=<n> levels=<l0>,<l1>,... batches=<b>"
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/sme.h"

int main(int argc, char** argv) {
  if (argc != 9 && argc != 10) {
    std::fprintf(stderr, "usage: %s graph.bin algo query cl ord vbegin vend budget [hub_min_degree]\n",
                 argv[0]);
    return 2;
  }
  uint64_t n = 0, nnz = 0;
  uint64_t* rp = nullptr;
  uint32_t* cols = nullptr;
  if (sme_bin_read(argv[1], &n, &nnz, &rp, &cols) != SME_OK) {
    std::fprintf(stderr, "read: %s\n", sme_last_error(nullptr));
    return 1;
  }
  const int algo = std::atoi(argv[2]), q = std::atoi(argv[3]), cl = std::atoi(argv[4]),
            ord = std::atoi(argv[5]);
  const uint32_t vb = static_cast<uint32_t>(std::strtoul(argv[6], nullptr, 10));
  uint64_t ve = std::strtoull(argv[7], nullptr, 10);
  if (ve > n) ve = n;
  const uint64_t budget = std::strtoull(argv[8], nullptr, 10);
  // SERVE="q:vb:ve,q:vb:ve,..." runs the requests through a one-slot server
  // (SERVE_MERGE=0: plain path) and prints "serve=<c1>,<c2>,..."
  if (const char* env = std::getenv("SERVE")) {
    std::vector<uint32_t> q, vbs, ves;
    for (const char* p = env; *p;) {
      char* e = nullptr;
      q.push_back(static_cast<uint32_t>(std::strtoul(p, &e, 10)));
      vbs.push_back(static_cast<uint32_t>(std::strtoul(e + 1, &e, 10)));
      ves.push_back(static_cast<uint32_t>(std::strtoul(e + 1, &e, 10)));
      p = *e == ',' ? e + 1 : e;
    }
    sme_server_config cfg{};
    cfg.mem_cap_bytes = budget;
    cfg.slots = 1;
    const char* m = std::getenv("SERVE_MERGE");
    cfg.merge = m ? std::atoi(m) : 1;
    cfg.max_merge = 64;
    cfg.min_reserve_bytes = budget / 4;
    sme_server* srv = nullptr;
    if (sme_server_open(0, rp, cols, n, nnz, &cfg, &srv) != SME_OK) {
      std::fprintf(stderr, "server: %s\n", sme_last_error(nullptr));
      return 1;
    }
    std::vector<uint64_t> ids(q.size());
    for (size_t i = 0; i < q.size(); ++i) {
      if (sme_server_submit(srv, static_cast<int>(q[i]), vbs[i], ves[i], &ids[i]) != SME_OK) {
        std::fprintf(stderr, "submit: %s\n", sme_last_error(nullptr));
        return 1;
      }
    }
    std::printf("serve=");
    for (size_t i = 0; i < q.size(); ++i) {
      uint64_t c = 0;
      double ms = 0;
      if (sme_server_wait(srv, ids[i], &c, &ms) != SME_OK) {
        std::fprintf(stderr, "wait: %s\n", sme_last_error(nullptr));
        return 1;
      }
      std::printf(i ? ",%llu" : "%llu", static_cast<unsigned long long>(c));
    }
    sme_server_stats st{};
    sme_server_get_stats(srv, &st);
    std::printf(" passes=%llu max_group=%llu\n", static_cast<unsigned long long>(st.passes),
                static_cast<unsigned long long>(st.max_group));
    sme_server_close(srv);
    sme_free_host(rp);
    sme_free_host(cols);
    return 0;
  }
  sme_ctx* ctx = nullptr;
  if (sme_open(0, budget, &ctx) != SME_OK) {
    std::fprintf(stderr, "open: %s\n", sme_last_error(nullptr));
    return 1;
  }
  if (argc == 10) sme_set_hub_bitmaps(ctx, std::strtoull(argv[9], nullptr, 10), 1ull << 30);
  // RANK=1 keeps a degree-ranked copy of the graph (clique plans over all
  // start vertices run on it; the output then says ranked=1)
  if (const char* env = std::getenv("RANK")) sme_set_rank_order(ctx, std::atoi(env));
  // TUNE="lane_path_max,item_grid,batch_bytes[,super_chunks[,stream_ratio[,group_max[,wg_path_min
  //       [,clique_level_engine[,census_level_engine]]]]]]"
  // sets the schedule tunables
  if (const char* env = std::getenv("TUNE")) {
    sme_tuning t{};
    char* s = nullptr;
    t.lane_path_max = std::strtoull(env, &s, 10);
    if (*s == ',') t.item_grid = static_cast<uint32_t>(std::strtoul(s + 1, &s, 10));
    if (*s == ',') t.batch_bytes = std::strtoull(s + 1, &s, 10);
    if (*s == ',') t.super_chunks = static_cast<uint32_t>(std::strtoul(s + 1, &s, 10));
    if (*s == ',') t.stream_ratio = static_cast<uint32_t>(std::strtoul(s + 1, &s, 10));
    if (*s == ',') t.group_max = static_cast<uint32_t>(std::strtoul(s + 1, &s, 10));
    if (*s == ',') t.wg_path_min = std::strtoull(s + 1, &s, 10);
    if (*s == ',') t.clique_level_engine = static_cast<uint32_t>(std::strtoul(s + 1, &s, 10));
    if (*s == ',') t.census_level_engine = static_cast<uint32_t>(std::strtoul(s + 1, &s, 10));
    sme_set_tuning(ctx, &t);
  }
  sme_plan* plan = nullptr;
  if (sme_plan_create(algo, q, cl, ord, &plan) != SME_OK) {
    std::fprintf(stderr, "plan: %s\n", sme_last_error(nullptr));
    return 1;
  }
  sme_graph* g = nullptr;
  if (sme_graph_upload(ctx, rp, cols, n, nnz, &g) != SME_OK) {
    std::fprintf(stderr, "upload: %s\n", sme_last_error(ctx));
    return 1;
  }
  // LIST=<cap> lists up to cap matches: "total=<t> rows=<r> list=<flattened rows>"
  if (const char* env = std::getenv("LIST")) {
    const uint64_t cap = std::strtoull(env, nullptr, 10);
    const int nv = sme_plan_vertices(plan);
    std::vector<uint32_t> out(cap * nv + 1);
    uint64_t total = 0, rows = 0;
    const int rc = sme_list_matches(ctx, g, plan, vb, static_cast<uint32_t>(ve), out.data(), cap, &total, &rows);
    if (rc != SME_OK) {
      std::fprintf(stderr, "list: rc=%d %s\n", rc, sme_last_error(ctx));
      return rc == SME_ERR_UNSUPPORTED ? 3 : 1;
    }
    std::printf("total=%llu rows=%llu list=", static_cast<unsigned long long>(total),
                static_cast<unsigned long long>(rows));
    for (uint64_t i = 0; i < rows * nv; ++i) std::printf(i ? ",%u" : "%u", out[i]);
    std::printf("\n");
    return 0;
  }
  uint64_t count = 0, levels[SME_MAX_LEVELS] = {0};
  sme_stats st{};
  // BUCKETS="b0,b1,...,bk" switches to sme_count_buckets over those bounds
  std::vector<uint32_t> bounds;
  if (const char* env = std::getenv("BUCKETS")) {
    for (const char* s = env; *s;) {
      bounds.push_back(static_cast<uint32_t>(std::strtoul(s, const_cast<char**>(&s), 10)));
      if (*s == ',') ++s;
    }
  }
  std::vector<uint64_t> bcounts(bounds.empty() ? 0 : bounds.size() - 1);
  // RANKED=1: the range is in the degree-ranked copy's ids (sme_count_ranked);
  // NOLEVELS=1: no level outputs are asked for
  const bool ranked_ids = std::getenv("RANKED") && std::atoi(std::getenv("RANKED"));
  // STATIC=1: the no-allocation entry point (sme_count_static)
  const bool no_alloc = std::getenv("STATIC") && std::atoi(std::getenv("STATIC"));
  uint64_t* lv = (std::getenv("NOLEVELS") && std::atoi(std::getenv("NOLEVELS"))) ? nullptr : levels;
  const int rc = !bounds.empty()
                     ? sme_count_buckets(ctx, g, plan, bounds.data(), bcounts.size(), bcounts.data(),
                                         &count, lv, &st)
                 : ranked_ids ? sme_count_ranked(ctx, g, plan, vb, static_cast<uint32_t>(ve), &count, lv, &st)
                 : no_alloc   ? sme_count_static(ctx, g, plan, vb, static_cast<uint32_t>(ve), &count, lv, &st)
                              : sme_count(ctx, g, plan, vb, static_cast<uint32_t>(ve), &count, lv, &st);
  if (rc != SME_OK) {
    std::fprintf(stderr, "count: rc=%d %s\n", rc, sme_last_error(ctx));
    return 1;
  }
  std::printf("count=%llu levels=", static_cast<unsigned long long>(count));
  for (int i = 0; i < sme_plan_levels(plan); ++i)
    std::printf(i ? ",%llu" : "%llu", static_cast<unsigned long long>(levels[i]));
  std::printf(" batches=%llu", static_cast<unsigned long long>(st.batches));
  if (st.ranked) std::printf(" ranked=1");
  if (!bounds.empty()) {
    std::printf(" buckets=");
    for (size_t k = 0; k < bcounts.size(); ++k)
      std::printf(k ? ",%llu" : "%llu", static_cast<unsigned long long>(bcounts[k]));
  }
  std::printf("\n");
  sme_graph_destroy(g);
  sme_plan_destroy(plan);
  sme_close(ctx);
  sme_free_host(rp);
  sme_free_host(cols);
  return 0;
}
