// cli.cpp -- patternmatchgpu: the command line of the reference GPU binary
// (src/gpu/PatternMatchGPU.cu:11-82) on top of the C ABI.
//
//   patternmatchgpu -f FILE(.bin|.txt) -d DIRECTED -a ALGO -q QUERY
//                   -m MODE -c COMPRESS_LEVEL -r ENABLE_ORDERING -v DEV_NUM
//                   -b BUDGET_GB
//
// Prints the reference's result line
//   total_match_count=<n>, elapsed_time=<ms>ms
// (src/gpu/executor/PatternMatch.cuh:85-87).  As in PatternMatch::Execute
// (:52-88) the timed region covers plan set-up, the CSR upload and the
// enumeration, not the file load.  Work is split into start-vertex tasks the
// way IntraPartScheduler::PrepareSchedule does for one partition
// (src/gpu/task/IntraPartScheduler.h:77-101).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sme.h"

namespace {

// CommandLine (src/utils/CommandLine.h:45-85): "-x value" pairs.
struct Args {
  int argc;
  char** argv;
  const char* Get(const char* opt) const {
    for (int i = 1; i + 1 < argc; ++i)
      if (std::strcmp(argv[i], opt) == 0) return argv[i + 1];
    return nullptr;
  }
  std::string Str(const char* opt, const std::string& def) const {
    const char* v = Get(opt);
    return v ? v : def;
  }
  long Int(const char* opt, long def) const {
    const char* v = Get(opt);
    return v ? std::atol(v) : def;
  }
};

void Usage() {
  std::printf(
      "./patternmatchgpu -f FILENAME -d IS_DIRECTED -a ALGORITHM -q QUERY -m EXECUTE_MODE "
      "-c COMPRESS_LEVEL -r ENABLE_ORDERING -v DEV_NUM -b DEVICE_BUDGET_GB\n"
      "-l MATERIALIZE (1: after the result line, print up to -L MAX_MATCHES matches, one per\n"
      "   line, data vertices of query vertices 0.., lexicographic; GPU_GPSM / GPU_GPSM_REUSE)\n"
      "ALGORITHM: GPU_GPSM=4, GPU_LIGHT=6\n"
      "EXECUTE_MODE: IN_MEMORY_GPU=1 (HYBRID_CPU_GPU=0 is accepted with one partition)\n"
      "QUERY: Q0 (triangle) 0, Q1 (square) 1, Q2 (chordal square) 2, Q3 (4 clique) 3, "
      "Q4 (house) 4, Q5 (quad triangle) 5, Q6 (near5clique) 6, Q7 (5 clique) 7, "
      "Q8 (chordal roof) 8, Q9 (three triangle) 9, Q10 (solar square) 10, Q11 (6 clique) 11, "
      "Q12 12, Q13 13\n");
}

int Die(const std::string& what, sme_ctx* ctx) {
  std::fprintf(stderr, "error: %s: %s\n", what.c_str(), sme_last_error(ctx));
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 1) {
    Usage();
    return 1;
  }
  Args a{argc, argv};
  const std::string filename = a.Str("-f", "");
  const int directed = static_cast<int>(a.Int("-d", 1));
  const int algo = static_cast<int>(a.Int("-a", SME_ALGO_LIGHT));
  const int query = static_cast<int>(a.Int("-q", 0));
  const int partition_num = static_cast<int>(a.Int("-p", 1));
  const int dev_num = static_cast<int>(a.Int("-v", 1));
  const int mode = static_cast<int>(a.Int("-m", 0));
  const int compress = static_cast<int>(a.Int("-c", SME_CL_SPECIAL));
  const int ordering = static_cast<int>(a.Int("-r", 1));
  const double budget_gb = std::atof(a.Str("-b", "8").c_str());
  const bool materialize = a.Int("-l", 0) != 0;  // parsed, never consumed by the reference
  const uint64_t max_matches = static_cast<uint64_t>(a.Int("-L", 1000000));
  std::printf("filename=%s,directed=%d,algo=%d,query_type=%d,execute_mode=%d,compress_level=%d,"
              "enable_ordering=%d,dev_num=%d,budget_gb=%.2f\n",
              filename.c_str(), directed, algo, query, mode, compress, ordering, dev_num, budget_gb);
  if (filename.empty()) {
    Usage();
    return 1;
  }
  if (partition_num != 1 || (mode != 0 && mode != 1)) {
    std::fprintf(stderr, "error: only the in-memory single-partition mode is built (-m 1 -p 1)\n");
    return 2;
  }
  if (dev_num != 1) {
    std::fprintf(stderr, "error: this build runs one GPU per process (-v 1)\n");
    return 2;
  }

  uint64_t n = 0, nnz = 0;
  uint64_t* rp = nullptr;
  uint32_t* cols = nullptr;
  const bool is_bin = filename.size() > 4 && filename.compare(filename.size() - 4, 4, ".bin") == 0;
  const int rs = is_bin ? sme_bin_read(filename.c_str(), &n, &nnz, &rp, &cols)
                        : sme_snap_read(filename.c_str(), directed, &n, &nnz, &rp, &cols);
  if (rs != SME_OK) return Die("reading " + filename, nullptr);
  std::printf("vertex_count=%llu,edge_count=%llu\n", static_cast<unsigned long long>(n),
              static_cast<unsigned long long>(nnz));

  sme_ctx* ctx = nullptr;
  if (sme_open(0, static_cast<uint64_t>(budget_gb * 1024.0 * 1024.0 * 1024.0), &ctx) != SME_OK)
    return Die("opening the device", nullptr);

  const auto t0 = std::chrono::steady_clock::now();
  sme_plan* plan = nullptr;
  if (sme_plan_create(algo, query, compress, ordering, &plan) != SME_OK) return Die("plan", ctx);
  sme_graph* g = nullptr;
  if (sme_graph_upload(ctx, rp, cols, n, nnz, &g) != SME_OK) return Die("graph upload", ctx);

  // two start-vertex tasks per device (IntraPartScheduler.h:77-101)
  const uint64_t units = 2;
  const uint64_t step = (n + units - 1) / units;
  unsigned long long total = 0;
  uint64_t batches = 0;
  for (uint64_t s = 0; s < n; s += step) {
    const uint64_t e = s + step < n ? s + step : n;
    uint64_t cnt = 0;
    sme_stats st{};
    if (sme_count(ctx, g, plan, static_cast<uint32_t>(s), static_cast<uint32_t>(e), &cnt, nullptr, &st) !=
        SME_OK)
      return Die("count", ctx);
    total += cnt;
    batches += st.batches;
  }
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::printf("batches=%llu\n", static_cast<unsigned long long>(batches));
  std::printf("total_match_count=%llu, elapsed_time=%.3fms\n", total, ms);

  if (materialize) {
    const int nv = sme_plan_vertices(plan);
    std::vector<uint32_t> rows(max_matches * static_cast<uint64_t>(nv) + 1);
    uint64_t listed_total = 0, written = 0;
    if (sme_list_matches(ctx, g, plan, 0, static_cast<uint32_t>(n), rows.data(), max_matches,
                         &listed_total, &written) != SME_OK)
      return Die("match list", ctx);
    for (uint64_t r = 0; r < written; ++r) {
      for (int w = 0; w < nv; ++w) std::printf(w ? " %u" : "%u", rows[r * nv + w]);
      std::printf("\n");
    }
  }

  sme_graph_destroy(g);
  sme_plan_destroy(plan);
  sme_close(ctx);
  sme_free_host(rp);
  sme_free_host(cols);
  return 0;
}
