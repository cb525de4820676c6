// serve.h -- pattern-count server over one resident graph.
//
// Requests (query, start-vertex range) are queued on arrival.  A dispatcher
// hands work to `slots` worker passes, each with its own context and stream
// on the one device, all carving one shared pool of mem_cap bytes:
//   - continuous batching: when a slot frees up, every pending request of
//     the oldest request's pattern (up to max_merge) becomes one pass over
//     the union of their ranges, counted per start-vertex bucket, so each
//     request's count is the sum of the buckets inside its range;
//   - admission: a pass starts only when its estimated intermediate bytes
//     (a per-pattern bytes-per-work-unit ratio learned from finished passes
//     times the summed squared degrees of the covered start vertices,
//     clamped to [min_reserve, mem_cap]; a pattern not seen yet gets
//     mem_cap / slots) fit beside the passes running; the pass's batch
//     planner then stays within that reservation.
// With merge off and one slot this is the plain path: one request at a time
// in arrival order, the whole budget each (IntraPartTask over the request's
// range, src/gpu/task/Task.h:23-43).
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"

namespace sme {

struct ServerConfig {
  uint64_t mem_cap = 0;
  int slots = 4;
  bool merge = true;
  uint64_t max_merge = 1024;
  uint64_t min_reserve = 1ull << 30;
  uint64_t hub_min_degree = 0;
  uint64_t hub_max_bytes = 0;
  Context::Tuning tuning;
};

struct ServerStats {
  uint64_t passes = 0;          // engine passes run
  uint64_t requests_done = 0;
  uint64_t max_group = 0;       // most requests served by one pass
  uint64_t peak_reserved = 0;   // largest sum of admitted reservations
  uint64_t peak_concurrent = 0; // most passes running at once
  uint64_t failed = 0;
};

class Server {
 public:
  ~Server();
  Status Open(int device, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n, uint64_t nnz,
              const ServerConfig& cfg, std::string* err);
  Status Submit(int query, uint32_t v_begin, uint32_t v_end, uint64_t* id, std::string* err);
  // Blocks until request `id` is done; latency is submit to completion.
  Status Wait(uint64_t id, uint64_t* count, double* latency_ms, std::string* err);
  ServerStats Stats();
  // Finishes every submitted request, then stops the threads.
  void Close();

 private:
  using Clock = std::chrono::steady_clock;
  struct Request {
    int q = 0;
    uint32_t vb = 0, ve = 0;
    Clock::time_point t_submit, t_done;
    uint64_t count = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    Status status = kOk;
    std::string err;
  };
  struct Job {
    int q = 0;
    std::vector<uint64_t> ids;
    uint64_t reserve = 0;
    uint64_t work = 0;
  };

  void Dispatch();
  void Work(int slot);
  uint64_t WorkUnits(uint32_t vb, uint32_t ve) const;
  // Runs one job on `slot`'s context; fills counts per request id and *peak
  // with the most pool bytes the job held at once (over all its passes).
  Status RunJob(int slot, const Job& job, const std::vector<Request>& snap,
                std::vector<uint64_t>* counts, std::string* err, uint64_t* peak);

  ServerConfig cfg_;
  DevicePool pool_;
  std::vector<std::unique_ptr<Context>> ctx_;
  std::unique_ptr<DevGraph> graph_;
  std::vector<EnginePlan> plans_;
  std::vector<uint64_t> work_prefix_;  // prefix sums of deg(v)^2
  std::vector<double> bytes_per_unit_; // learned per query; < 0: not yet seen

  std::mutex mu_;
  std::condition_variable cv_dispatch_, cv_done_;
  std::vector<std::unique_ptr<std::condition_variable>> cv_slot_;
  std::deque<uint64_t> queue_;
  std::deque<Request> reqs_;  // index = id; deque keeps references stable
  std::vector<std::unique_ptr<Job>> slot_job_;
  std::vector<bool> slot_busy_;
  uint64_t reserved_ = 0;
  uint64_t running_ = 0;
  bool closing_ = false;
  bool stop_ = false;
  bool open_ = false;
  ServerStats stats_;
  std::thread dispatcher_;
  std::vector<std::thread> workers_;
};

}  // namespace sme
