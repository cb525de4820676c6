// capi.cpp -- the extern "C" boundary declared in include/sme.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sme.h"
#include "engine.h"
#include "graph.h"
#include "query.h"
#include "serve.h"

struct sme_ctx {
  sme::Context impl;
};
struct sme_graph {
  sme::DevGraph impl;
};
struct sme_plan {
  sme::EnginePlan plan;
  std::string text;
};
struct sme_server {
  sme::Server impl;
};

namespace {

thread_local std::string g_last_error;  // errors without a context

int Report(sme_ctx* ctx, int status, const std::string& msg) {
  if (status != SME_OK) {
    if (ctx) ctx->impl.last_error = msg;
    g_last_error = msg;
  }
  return status;
}

// Runs an entry point's body so that no C++ exception crosses the C ABI: an
// engine error keeps its status, host allocation failure is SME_ERR_OOM,
// anything else SME_ERR_INVALID.
template <typename F>
int Guarded(sme_ctx* ctx, F&& body) {
  try {
    return body();
  } catch (const sme::Error& e) {
    return Report(ctx, e.status, e.msg);
  } catch (const std::bad_alloc&) {
    return Report(ctx, SME_ERR_OOM, "host memory exhausted");
  } catch (const std::exception& e) {
    return Report(ctx, SME_ERR_INVALID, e.what());
  } catch (...) {
    return Report(ctx, SME_ERR_INVALID, "unexpected failure");
  }
}

template <typename T>
T* CopyOut(const std::vector<T>& v) {
  T* p = static_cast<T*>(std::malloc(v.empty() ? sizeof(T) : v.size() * sizeof(T)));
  if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

int HandOut(const sme::HostGraph& g, uint64_t* n, uint64_t* nnz, uint64_t** rp, uint32_t** cols) {
  *n = g.n;
  *nnz = g.nnz;
  *rp = CopyOut(g.row_ptrs);
  *cols = CopyOut(g.cols);
  if (!*rp || !*cols) {
    std::free(*rp);
    std::free(*cols);
    *rp = nullptr;
    *cols = nullptr;
    return Report(nullptr, SME_ERR_OOM, "host allocation failed");
  }
  return SME_OK;
}

}  // namespace

extern "C" {

int sme_device_count(int* out) {
  if (!out) return SME_ERR_INVALID;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    c = 0;
  }
  *out = c;
  return SME_OK;
}

int sme_open(int device, uint64_t mem_budget_bytes, sme_ctx** out) {
  return Guarded(nullptr, [&]() -> int {
    if (!out) return Report(nullptr, SME_ERR_INVALID, "null output pointer");
    *out = nullptr;
    sme_ctx* c = new (std::nothrow) sme_ctx;
    if (!c) return Report(nullptr, SME_ERR_OOM, "host allocation failed");
    std::string err;
    const sme::Status s = c->impl.Open(device, mem_budget_bytes, &err);
    if (s != sme::kOk) {
      delete c;
      return Report(nullptr, s, err);
    }
    *out = c;
    return SME_OK;
  });
}

void sme_close(sme_ctx* ctx) { delete ctx; }

const char* sme_last_error(const sme_ctx* ctx) {
  if (ctx && !ctx->impl.last_error.empty()) return ctx->impl.last_error.c_str();
  return g_last_error.c_str();
}

int sme_graph_upload(sme_ctx* ctx, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                     uint64_t nnz, sme_graph** out) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !out) return Report(ctx, SME_ERR_INVALID, "null argument");
    *out = nullptr;
    sme_graph* g = new (std::nothrow) sme_graph;
    if (!g) return Report(ctx, SME_ERR_OOM, "host allocation failed");
    std::string err;
    sme::Status s;
    try {
      s = g->impl.Upload(&ctx->impl, row_ptrs, cols, n, nnz, &err);
    } catch (const sme::Error& e) {
      s = e.status;
      err = e.msg;
    } catch (...) {
      delete g;
      throw;
    }
    if (s != sme::kOk) {
      delete g;
      return Report(ctx, s, err);
    }
    *out = g;
    return SME_OK;
  });
}

void sme_graph_destroy(sme_graph* g) { delete g; }

int sme_plan_create(int algo, int query_type, int compress_level, int enable_ordering,
                    sme_plan** out) {
  return Guarded(nullptr, [&]() -> int {
    if (!out) return Report(nullptr, SME_ERR_INVALID, "null output pointer");
    *out = nullptr;
    sme_plan* p = new (std::nothrow) sme_plan;
    if (!p) return Report(nullptr, SME_ERR_OOM, "host allocation failed");
    bool ok = false;
    if (algo == SME_ALGO_LIGHT) {
      ok = sme::MakeLightEnginePlan(query_type, compress_level, enable_ordering != 0, &p->plan, &p->text);
    } else if (algo == SME_ALGO_GPSM || algo == SME_ALGO_GPSM_REUSE) {
      ok = sme::MakeGpsmEnginePlan(query_type, enable_ordering != 0, algo == SME_ALGO_GPSM_REUSE,
                                   &p->plan, &p->text);
    } else {
      delete p;
      return Report(nullptr, SME_ERR_UNSUPPORTED, "algorithm " + std::to_string(algo) + " is not supported");
    }
    if (!ok) {
      delete p;
      return Report(nullptr, SME_ERR_INVALID,
                    "cannot build a plan for query " + std::to_string(query_type) + " (compress level " +
                        std::to_string(compress_level) + ")");
    }
    *out = p;
    return SME_OK;
  });
}

int sme_plan_describe(const sme_plan* p, char* buf, size_t cap) {
  return Guarded(nullptr, [&]() -> int {
    if (!p) return -1;
    if (buf && cap) {
      const size_t n = p->text.size() < cap - 1 ? p->text.size() : cap - 1;
      std::memcpy(buf, p->text.data(), n);
      buf[n] = '\0';
    }
    return static_cast<int>(p->text.size());
  });
}

int sme_plan_levels(const sme_plan* p) { return p ? p->plan.nops : -1; }

int sme_plan_set_reuse(sme_plan* p, int on) {
  if (!p) return SME_ERR_INVALID;
  p->plan.reuse = on ? 1 : 0;
  return SME_OK;
}

void sme_plan_destroy(sme_plan* p) { delete p; }

int sme_count(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t v_begin, uint32_t v_end,
              uint64_t* out_count, uint64_t* level_counts, sme_stats* stats) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !g || !p || !out_count) return Report(ctx, SME_ERR_INVALID, "null argument");
    std::string err;
    sme::RunStats rs;
    const sme::Status s =
        sme::RunPlan(&ctx->impl, &g->impl, &p->plan, v_begin, v_end, out_count, level_counts, &rs, &err);
    if (s != sme::kOk) return Report(ctx, s, err);
    if (stats) {
      stats->batches = rs.batches;
      stats->d2h_syncs = rs.d2h_syncs;
      stats->peak_pool_bytes = rs.peak_pool_bytes;
      stats->elapsed_ms = rs.elapsed_ms;
      stats->ranked = rs.ranked;
    }
    return SME_OK;
  });
}

int sme_count_static(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t v_begin,
                     uint32_t v_end, uint64_t* out_count, uint64_t* level_counts, sme_stats* stats) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !g || !p || !out_count) return Report(ctx, SME_ERR_INVALID, "null argument");
    std::string err;
    sme::RunStats rs;
    const sme::Status s = sme::RunPlanNoAlloc(&ctx->impl, &g->impl, &p->plan, v_begin, v_end, out_count,
                                              level_counts, &rs, &err);
    if (s != sme::kOk) return Report(ctx, s, err);
    if (stats) {
      stats->batches = rs.batches;
      stats->d2h_syncs = rs.d2h_syncs;
      stats->peak_pool_bytes = rs.peak_pool_bytes;
      stats->elapsed_ms = rs.elapsed_ms;
      stats->ranked = rs.ranked;
    }
    return SME_OK;
  });
}

uint64_t sme_static_scratch_bytes(void) { return sme::StaticScratchBytes(); }

int sme_count_ranked(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t r_begin,
                     uint32_t r_end, uint64_t* out_count, uint64_t* level_counts, sme_stats* stats) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !g || !p || !out_count) return Report(ctx, SME_ERR_INVALID, "null argument");
    std::string err;
    sme::RunStats rs;
    const sme::Status s = sme::RunPlan(&ctx->impl, &g->impl, &p->plan, r_begin, r_end, out_count,
                                       level_counts, &rs, &err, true);
    if (s != sme::kOk) return Report(ctx, s, err);
    if (stats) {
      stats->batches = rs.batches;
      stats->d2h_syncs = rs.d2h_syncs;
      stats->peak_pool_bytes = rs.peak_pool_bytes;
      stats->elapsed_ms = rs.elapsed_ms;
      stats->ranked = rs.ranked;
    }
    return SME_OK;
  });
}

int sme_count_buckets(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, const uint32_t* bounds,
                      uint64_t nbuckets, uint64_t* counts, uint64_t* out_total, uint64_t* level_counts,
                      sme_stats* stats) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !g || !p || !bounds || !counts || !out_total)
      return Report(ctx, SME_ERR_INVALID, "null argument");
    std::string err;
    sme::RunStats rs;
    const sme::Status s = sme::RunPlanBuckets(&ctx->impl, &g->impl, &p->plan, bounds, nbuckets, counts,
                                              out_total, level_counts, &rs, &err);
    if (s != sme::kOk) return Report(ctx, s, err);
    if (stats) {
      stats->batches = rs.batches;
      stats->d2h_syncs = rs.d2h_syncs;
      stats->peak_pool_bytes = rs.peak_pool_bytes;
      stats->elapsed_ms = rs.elapsed_ms;
      stats->ranked = rs.ranked;
    }
    return SME_OK;
  });
}

int sme_list_matches(sme_ctx* ctx, const sme_graph* g, const sme_plan* p, uint32_t v_begin,
                     uint32_t v_end, uint32_t* out, uint64_t max_matches, uint64_t* total,
                     uint64_t* written) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !g || !p || !total || !written) return Report(ctx, SME_ERR_INVALID, "null argument");
    std::string err;
    const sme::Status s = sme::ListMatches(&ctx->impl, &g->impl, &p->plan, v_begin, v_end, out,
                                           max_matches, total, written, &err);
    return s == sme::kOk ? SME_OK : Report(ctx, s, err);
  });
}

int sme_plan_vertices(const sme_plan* p) { return p ? p->plan.nv : -1; }

int sme_set_tuning(sme_ctx* ctx, const sme_tuning* t) {
  if (!ctx || !t) return Report(ctx, SME_ERR_INVALID, "null argument");
  ctx->impl.tuning.lane_path_max = t->lane_path_max;
  ctx->impl.tuning.item_grid = t->item_grid;
  ctx->impl.tuning.super_chunks = t->super_chunks;
  ctx->impl.tuning.batch_bytes = t->batch_bytes;
  ctx->impl.tuning.stream_ratio = t->stream_ratio;
  ctx->impl.tuning.group_max = t->group_max;
  ctx->impl.tuning.wg_path_min = t->wg_path_min;
  ctx->impl.tuning.clique_level_engine = t->clique_level_engine;
  ctx->impl.tuning.census_level_engine = t->census_level_engine;
  return SME_OK;
}

int sme_get_tuning(const sme_ctx* ctx, sme_tuning* t) {
  if (!ctx || !t) return SME_ERR_INVALID;
  *t = sme_tuning{};
  t->lane_path_max = ctx->impl.tuning.lane_path_max;
  t->item_grid = ctx->impl.tuning.item_grid;
  t->super_chunks = ctx->impl.tuning.super_chunks;
  t->batch_bytes = ctx->impl.tuning.batch_bytes;
  t->stream_ratio = ctx->impl.tuning.stream_ratio;
  t->group_max = ctx->impl.tuning.group_max;
  t->wg_path_min = ctx->impl.tuning.wg_path_min;
  t->clique_level_engine = ctx->impl.tuning.clique_level_engine;
  t->census_level_engine = ctx->impl.tuning.census_level_engine;
  return SME_OK;
}

int sme_autotune(sme_ctx* ctx, const sme_graph* g, const sme_plan* const* plans, int nplans,
                 const uint32_t* bounds, uint64_t nbuckets, double sample_fraction,
                 sme_tuning* best, double* best_ms, sme_tune_trial* trials, int cap,
                 int* ntrials) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !g || !plans || nplans <= 0 || !best)
      return Report(ctx, SME_ERR_INVALID, "null argument");
    if (!(sample_fraction > 0.0) || sample_fraction > 1.0)
      return Report(ctx, SME_ERR_INVALID, "sample_fraction must be in (0, 1]");
    for (int i = 0; i < nplans; ++i)
      if (!plans[i]) return Report(ctx, SME_ERR_INVALID, "null plan");
    // the sample: 8 evenly spaced blocks of buckets or of start vertices
    constexpr int kBlocks = 8;
    std::vector<std::pair<uint64_t, uint64_t>> blocks;  // [lo, hi) of buckets or vertices
    const uint64_t units = bounds ? nbuckets : g->impl.n;
    if (units == 0) return Report(ctx, SME_ERR_INVALID, "empty sample space");
    const uint64_t per = std::max<uint64_t>(
        1, static_cast<uint64_t>(sample_fraction * static_cast<double>(units) / kBlocks + 0.5));
    for (int b = 0; b < kBlocks; ++b) {
      const uint64_t lo = units * b / kBlocks;
      const uint64_t hi = std::min(units, lo + per);
      if (hi > lo && (blocks.empty() || lo >= blocks.back().second)) blocks.push_back({lo, hi});
    }
    sme::Context& c = ctx->impl;
    std::string err;
    std::vector<uint64_t> counts;
    // one setting: the sample runs kRepeats times, the median total counts
    constexpr int kRepeats = 3;
    auto measure = [&](const sme::Context::Tuning& t, double* ms) -> int {
      c.tuning = t;
      double reps[kRepeats];
      for (int rep = 0; rep < kRepeats; ++rep) {
        double total = 0;
        for (int i = 0; i < nplans; ++i) {
          for (const auto& blk : blocks) {
            uint64_t cnt = 0;
            sme::RunStats rs;
            sme::Status s;
            if (bounds) {
              counts.assign(blk.second - blk.first, 0);
              s = sme::RunPlanBuckets(&c, &g->impl, &plans[i]->plan, bounds + blk.first,
                                      blk.second - blk.first, counts.data(), &cnt, nullptr, &rs, &err);
            } else {
              s = sme::RunPlan(&c, &g->impl, &plans[i]->plan, static_cast<uint32_t>(blk.first),
                               static_cast<uint32_t>(blk.second), &cnt, nullptr, &rs, &err);
            }
            if (s != sme::kOk) return Report(ctx, s, err);
            total += rs.elapsed_ms;
          }
        }
        reps[rep] = total;
      }
      std::sort(reps, reps + kRepeats);
      *ms = reps[kRepeats / 2];
      return SME_OK;
    };
    auto to_c = [](const sme::Context::Tuning& t) {
      return sme_tuning{t.lane_path_max, t.item_grid, t.super_chunks, t.batch_bytes,
                        t.stream_ratio, t.group_max, t.wg_path_min, t.clique_level_engine,
                        t.census_level_engine};
    };
    auto same = [](const sme::Context::Tuning& a, const sme::Context::Tuning& b) {
      return a.lane_path_max == b.lane_path_max && a.item_grid == b.item_grid &&
             a.super_chunks == b.super_chunks && a.batch_bytes == b.batch_bytes &&
             a.stream_ratio == b.stream_ratio && a.group_max == b.group_max &&
             a.wg_path_min == b.wg_path_min && a.clique_level_engine == b.clique_level_engine &&
             a.census_level_engine == b.census_level_engine;
    };
    const sme::Context::Tuning start = c.tuning;
    int n = 0;
    auto record = [&](const sme::Context::Tuning& t, double ms) {
      if (trials && n < cap) {
        trials[n].tuning = to_c(t);
        trials[n].ms = ms;
      }
      ++n;
    };
    // warm the caches and the pool once, then coordinate descent
    double ms = 0;
    int rc = measure(start, &ms);
    if (rc != SME_OK) {
      c.tuning = start;
      return rc;
    }
    sme::Context::Tuning cur = start;
    double cur_ms = 0;
    rc = measure(cur, &cur_ms);
    if (rc != SME_OK) {
      c.tuning = start;
      return rc;
    }
    record(cur, cur_ms);
    // a setting replaces the current one only when its median is at least
    // 2 % faster, so near-ties keep the earlier (simpler) choice
    constexpr double kGain = 0.98;
    const uint64_t lane_opts[] = {0, 8, 32, 128, 512};
    const uint32_t super_opts[] = {0, 2, 4, 8, 16};
    const uint32_t grid_opts[] = {0, 8192, 4096, 2048, 1024};
    const uint64_t batch_opts[] = {0, 4ull << 30, 1ull << 30, 256ull << 20, 64ull << 20};
    const uint32_t stream_opts[] = {0, 4, 8, 16, 32};
    const uint32_t group_opts[] = {0, 128, 512};
    const uint64_t wg_opts[] = {0, 2048, 8192};
    const int nopts[] = {5, 5, 5, 5, 5, 3, 3};
    for (int dim = 0; dim < 7; ++dim) {
      sme::Context::Tuning dim_best = cur;
      double dim_ms = cur_ms;
      for (int k = 0; k < nopts[dim]; ++k) {
        sme::Context::Tuning t = cur;
        if (dim == 0) t.lane_path_max = lane_opts[k];
        if (dim == 1) t.super_chunks = super_opts[k];
        if (dim == 2) t.item_grid = grid_opts[k];
        if (dim == 3) t.batch_bytes = batch_opts[k];
        if (dim == 4) t.stream_ratio = stream_opts[k];
        if (dim == 5) t.group_max = group_opts[k];
        if (dim == 6) t.wg_path_min = wg_opts[k];
        if (same(t, cur)) continue;
        double t_ms = 0;
        rc = measure(t, &t_ms);
        if (rc != SME_OK) {
          c.tuning = start;
          return rc;
        }
        record(t, t_ms);
        if (t_ms < dim_ms * kGain) {
          dim_best = t;
          dim_ms = t_ms;
        }
      }
      cur = dim_best;
      cur_ms = dim_ms;
    }
    c.tuning = cur;
    *best = to_c(cur);
    if (best_ms) *best_ms = cur_ms;
    if (ntrials) *ntrials = n;
    return SME_OK;
  });
}

int sme_server_open(int device, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                    uint64_t nnz, const sme_server_config* cfg, sme_server** out) {
  return Guarded(nullptr, [&]() -> int {
    if (!cfg || !out) return Report(nullptr, SME_ERR_INVALID, "null argument");
    *out = nullptr;
    sme_server* s = new (std::nothrow) sme_server;
    if (!s) return Report(nullptr, SME_ERR_OOM, "host allocation failed");
    sme::ServerConfig c;
    c.mem_cap = cfg->mem_cap_bytes;
    c.slots = cfg->slots;
    c.merge = cfg->merge != 0;
    c.max_merge = cfg->max_merge;
    c.min_reserve = cfg->min_reserve_bytes;
    c.hub_min_degree = cfg->hub_min_degree;
    c.hub_max_bytes = cfg->hub_max_bytes;
    c.tuning.lane_path_max = cfg->tuning.lane_path_max;
    c.tuning.item_grid = cfg->tuning.item_grid;
    c.tuning.super_chunks = cfg->tuning.super_chunks;
    c.tuning.batch_bytes = cfg->tuning.batch_bytes;
    c.tuning.stream_ratio = cfg->tuning.stream_ratio;
    c.tuning.group_max = cfg->tuning.group_max;
    c.tuning.wg_path_min = cfg->tuning.wg_path_min;
    c.tuning.clique_level_engine = cfg->tuning.clique_level_engine;
    c.tuning.census_level_engine = cfg->tuning.census_level_engine;
    std::string err;
    const sme::Status st = s->impl.Open(device, row_ptrs, cols, n, nnz, c, &err);
    if (st != sme::kOk) {
      delete s;
      return Report(nullptr, st, err);
    }
    *out = s;
    return SME_OK;
  });
}

int sme_server_submit(sme_server* s, int query, uint32_t v_begin, uint32_t v_end, uint64_t* id) {
  return Guarded(nullptr, [&]() -> int {
    if (!s || !id) return Report(nullptr, SME_ERR_INVALID, "null argument");
    std::string err;
    const sme::Status st = s->impl.Submit(query, v_begin, v_end, id, &err);
    return st == sme::kOk ? SME_OK : Report(nullptr, st, err);
  });
}

int sme_server_wait(sme_server* s, uint64_t id, uint64_t* count, double* latency_ms) {
  return Guarded(nullptr, [&]() -> int {
    if (!s) return Report(nullptr, SME_ERR_INVALID, "null argument");
    std::string err;
    const sme::Status st = s->impl.Wait(id, count, latency_ms, &err);
    return st == sme::kOk ? SME_OK : Report(nullptr, st, err);
  });
}

int sme_server_get_stats(sme_server* s, sme_server_stats* out) {
  return Guarded(nullptr, [&]() -> int {
    if (!s || !out) return SME_ERR_INVALID;
    const sme::ServerStats st = s->impl.Stats();
    out->passes = st.passes;
    out->requests_done = st.requests_done;
    out->max_group = st.max_group;
    out->peak_reserved = st.peak_reserved;
    out->peak_concurrent = st.peak_concurrent;
    out->failed = st.failed;
    return SME_OK;
  });
}

void sme_server_close(sme_server* s) { delete s; }

int sme_set_hub_bitmaps(sme_ctx* ctx, uint64_t min_degree, uint64_t max_bytes) {
  if (!ctx) return SME_ERR_INVALID;
  ctx->impl.hub_min_degree = min_degree;
  ctx->impl.hub_max_bytes = max_bytes;
  return SME_OK;
}

int sme_graph_hubs(const sme_graph* g, uint64_t* nhub) {
  if (!g || !nhub) return SME_ERR_INVALID;
  *nhub = g->impl.nhub;
  return SME_OK;
}

int sme_set_rank_order(sme_ctx* ctx, int on) {
  if (!ctx) return SME_ERR_INVALID;
  ctx->impl.rank_order = on != 0;
  return SME_OK;
}

int sme_graph_ranked(const sme_graph* g, int* on) {
  if (!g || !on) return SME_ERR_INVALID;
  *on = g->impl.ranked ? 1 : 0;
  return SME_OK;
}

int sme_profile_enable(sme_ctx* ctx, int on) {
  if (!ctx) return SME_ERR_INVALID;
  ctx->impl.profiling = on != 0;
  return SME_OK;
}

int sme_profile_reset(sme_ctx* ctx) {
  if (!ctx) return SME_ERR_INVALID;
  ctx->impl.ResetProfile();
  return SME_OK;
}

int sme_profile_slots(void) { return sme::kNumSlots; }

int sme_profile_get(const sme_ctx* ctx, int slot, char* name, size_t cap, uint64_t* launches, double* ms,
                    double* bytes, uint64_t* items) {
  if (!ctx || slot < 0 || slot >= sme::kNumSlots) return SME_ERR_INVALID;
  const sme::SlotStats& s = ctx->impl.slots[slot];
  if (name && cap) {
    std::strncpy(name, sme::SlotName(slot), cap - 1);
    name[cap - 1] = '\0';
  }
  if (launches) *launches = s.launches;
  if (ms) *ms = s.ms;
  if (bytes) *bytes = s.bytes;
  if (items) *items = s.items;
  return SME_OK;
}

int sme_calibrate_copy(sme_ctx* ctx, uint64_t bytes, double* ms) {
  return Guarded(ctx, [&]() -> int {
    if (!ctx || !ms || bytes == 0) return Report(ctx, SME_ERR_INVALID, "bad argument");
    sme::DevicePool& pool = ctx->impl.pool();
    void* a = pool.Alloc(bytes);
    void* b = a ? pool.Alloc(bytes) : nullptr;
    if (!a || !b) {
      pool.Free(a);
      return Report(ctx, SME_ERR_OOM, "pool too small for the copy");
    }
    hipStream_t st = ctx->impl.stream();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SME_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemsetAsync(a, 1, bytes, st) != hipSuccess || hipEventRecord(e0, st) != hipSuccess ||
        hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipEventRecord(e1, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
      rc = SME_ERR_HIP;
    } else {
      float f = 0;
      (void)hipEventElapsedTime(&f, e0, e1);
      *ms = f;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pool.Free(b);
    pool.Free(a);
    return rc == SME_OK ? SME_OK : Report(ctx, rc, "copy failed");
  });
}

int sme_bin_read(const char* path, uint64_t* n, uint64_t* nnz, uint64_t** row_ptrs, uint32_t** cols) {
  return Guarded(nullptr, [&]() -> int {
    if (!path || !n || !nnz || !row_ptrs || !cols) return Report(nullptr, SME_ERR_INVALID, "null argument");
    sme::HostGraph g;
    std::string err;
    const sme::Status s = sme::ReadBin(path, &g, &err);
    if (s != sme::kOk) return Report(nullptr, s, err);
    return HandOut(g, n, nnz, row_ptrs, cols);
  });
}

int sme_bin_write(const char* path, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                  uint64_t nnz) {
  return Guarded(nullptr, [&]() -> int {
    if (!path || !row_ptrs || (nnz && !cols)) return Report(nullptr, SME_ERR_INVALID, "null argument");
    sme::HostGraph g;
    g.n = n;
    g.nnz = nnz;
    g.row_ptrs.assign(row_ptrs, row_ptrs + n + 1);
    g.cols.assign(cols, cols + nnz);
    std::string err;
    const sme::Status s = sme::WriteBin(path, g, &err);
    return Report(nullptr, s, err);
  });
}

int sme_snap_read(const char* path, int directed, uint64_t* n, uint64_t* nnz, uint64_t** row_ptrs,
                  uint32_t** cols) {
  return Guarded(nullptr, [&]() -> int {
    if (!path || !n || !nnz || !row_ptrs || !cols) return Report(nullptr, SME_ERR_INVALID, "null argument");
    sme::HostGraph g;
    std::string err;
    const sme::Status s = sme::ReadSnap(path, directed != 0, &g, &err);
    if (s != sme::kOk) return Report(nullptr, s, err);
    return HandOut(g, n, nnz, row_ptrs, cols);
  });
}

int sme_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, uint64_t* n,
                 uint64_t* nnz, uint64_t** row_ptrs, uint32_t** cols) {
  return Guarded(nullptr, [&]() -> int {
    if (scale < 1 || scale > 31 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
      return Report(nullptr, SME_ERR_INVALID, "bad R-MAT parameters");
    if (!n || !nnz || !row_ptrs || !cols) return Report(nullptr, SME_ERR_INVALID, "null argument");
    sme::HostGraph g;
    sme::GenerateRmat(scale, edge_factor, a, b, c, seed, &g);
    return HandOut(g, n, nnz, row_ptrs, cols);
  });
}

int sme_gen_er(uint64_t n_vertices, uint64_t n_edges, uint64_t seed, uint64_t* n, uint64_t* nnz,
               uint64_t** row_ptrs, uint32_t** cols) {
  return Guarded(nullptr, [&]() -> int {
    if (n_vertices < 2 || n_vertices >= 0xffffffffull) return Report(nullptr, SME_ERR_INVALID, "bad vertex count");
    if (!n || !nnz || !row_ptrs || !cols) return Report(nullptr, SME_ERR_INVALID, "null argument");
    sme::HostGraph g;
    sme::GenerateErdosRenyi(n_vertices, n_edges, seed, &g);
    return HandOut(g, n, nnz, row_ptrs, cols);
  });
}

int sme_gen_er_batch(uint64_t ngraphs, uint64_t n_min, uint64_t n_max, double deg_min,
                     double deg_max, uint64_t seed, uint64_t* n, uint64_t* nnz,
                     uint64_t** row_ptrs, uint32_t** cols, uint64_t** graph_offs) {
  return Guarded(nullptr, [&]() -> int {
    if (!n || !nnz || !row_ptrs || !cols || !graph_offs)
      return Report(nullptr, SME_ERR_INVALID, "null argument");
    if (ngraphs == 0 || n_min < 2 || n_max < n_min || !(deg_min >= 0.0) || deg_max < deg_min)
      return Report(nullptr, SME_ERR_INVALID, "bad member ranges");
    if (ngraphs > (0xffffffffull - 1) / n_max)
      return Report(nullptr, SME_ERR_INVALID, "packed vertex ids exceed 32 bits");
    sme::HostGraph g;
    std::vector<uint64_t> offs;
    sme::GenerateErdosRenyiBatch(ngraphs, n_min, n_max, deg_min, deg_max, seed, &g, &offs);
    *graph_offs = CopyOut(offs);
    if (!*graph_offs) return Report(nullptr, SME_ERR_OOM, "host allocation failed");
    const int rc = HandOut(g, n, nnz, row_ptrs, cols);
    if (rc != SME_OK) {
      std::free(*graph_offs);
      *graph_offs = nullptr;
    }
    return rc;
  });
}

void sme_free_host(void* p) { std::free(p); }

}  // extern "C"
