// serve.cpp -- see serve.h.
#include "serve.h"

#include <algorithm>
#include <limits>

namespace sme {

Server::~Server() { Close(); }

Status Server::Open(int device, const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n,
                    uint64_t nnz, const ServerConfig& cfg, std::string* err) {
  if (open_) {
    *err = "server already open";
    return kErrInvalid;
  }
  if (cfg.slots < 1 || cfg.slots > 64) {
    *err = "slots must be in [1, 64]";
    return kErrInvalid;
  }
  if (cfg.mem_cap < (64ull << 20)) {
    *err = "mem_cap must be at least 64 MiB";
    return kErrInvalid;
  }
  cfg_ = cfg;
  if (cfg_.min_reserve > cfg_.mem_cap) cfg_.min_reserve = cfg_.mem_cap;
  if (cfg_.max_merge == 0) cfg_.max_merge = 1;
  for (int i = 0; i < cfg_.slots; ++i) {
    ctx_.emplace_back(new Context);
    const Status s = ctx_.back()->OpenShared(device, &pool_, err);
    if (s != kOk) return s;
    ctx_.back()->tuning = cfg_.tuning;
  }
  if (!pool_.Init(cfg_.mem_cap)) {
    *err = "cannot reserve the server pool of " + std::to_string(cfg_.mem_cap) + " bytes";
    return kErrOom;
  }
  pool_.SetShared(cfg_.slots > 1);
  ctx_[0]->hub_min_degree = cfg_.hub_min_degree;
  ctx_[0]->hub_max_bytes = cfg_.hub_max_bytes;
  graph_.reset(new DevGraph);
  try {
    const Status s = graph_->Upload(ctx_[0].get(), row_ptrs, cols, n, nnz, err);
    if (s != kOk) return s;
  } catch (const Error& e) {
    *err = e.msg;
    return e.status;
  }
  // LIGHT plans of every query, SPECIAL compress level
  plans_.resize(kNumQueries);
  for (int q = 0; q < kNumQueries; ++q) {
    std::string text;
    if (!MakeLightEnginePlan(q, 3, true, &plans_[q], &text)) {
      *err = "cannot build the plan of query " + std::to_string(q);
      return kErrInvalid;
    }
  }
  // work units: squared degree, the wedge count a start vertex roots
  work_prefix_.assign(n + 1, 0);
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t d = row_ptrs[v + 1] - row_ptrs[v];
    work_prefix_[v + 1] = work_prefix_[v] + d * d + 1;
  }
  bytes_per_unit_.assign(kNumQueries, -1.0);
  cv_slot_.clear();
  for (int i = 0; i < cfg_.slots; ++i) cv_slot_.emplace_back(new std::condition_variable);
  slot_job_.resize(cfg_.slots);
  slot_busy_.assign(cfg_.slots, false);
  open_ = true;
  dispatcher_ = std::thread([this] { Dispatch(); });
  for (int i = 0; i < cfg_.slots; ++i) workers_.emplace_back([this, i] { Work(i); });
  return kOk;
}

uint64_t Server::WorkUnits(uint32_t vb, uint32_t ve) const {
  return ve > vb ? work_prefix_[ve] - work_prefix_[vb] : 0;
}

Status Server::Submit(int query, uint32_t v_begin, uint32_t v_end, uint64_t* id, std::string* err) {
  if (!open_) {
    *err = "server is not open";
    return kErrInvalid;
  }
  if (query < 0 || query >= kNumQueries) {
    *err = "unknown query " + std::to_string(query);
    return kErrInvalid;
  }
  const uint64_t n = graph_->n;
  if (v_end > n) v_end = static_cast<uint32_t>(n);
  std::lock_guard<std::mutex> lk(mu_);
  if (closing_) {
    *err = "server is closing";
    return kErrInvalid;
  }
  Request r;
  r.q = query;
  r.vb = v_begin;
  r.ve = v_end;
  r.t_submit = Clock::now();
  *id = reqs_.size();
  if (v_end <= v_begin) {
    r.state = 2;
    r.t_done = r.t_submit;
    reqs_.push_back(r);
    stats_.requests_done++;
    cv_done_.notify_all();
    return kOk;
  }
  reqs_.push_back(r);
  queue_.push_back(*id);
  cv_dispatch_.notify_one();
  return kOk;
}

Status Server::Wait(uint64_t id, uint64_t* count, double* latency_ms, std::string* err) {
  std::unique_lock<std::mutex> lk(mu_);
  if (id >= reqs_.size()) {
    *err = "unknown request id";
    return kErrInvalid;
  }
  cv_done_.wait(lk, [&] { return reqs_[id].state >= 2; });
  const Request& r = reqs_[id];
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  if (r.state == 3) {
    *err = r.err;
    return r.status;
  }
  if (count) *count = r.count;
  return kOk;
}

ServerStats Server::Stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

void Server::Dispatch() {
  std::unique_lock<std::mutex> lk(mu_);
  while (true) {
    auto free_slot = [&]() {
      for (int i = 0; i < cfg_.slots; ++i)
        if (!slot_busy_[i]) return i;
      return -1;
    };
    cv_dispatch_.wait(lk, [&] {
      return stop_ || (!queue_.empty() && free_slot() >= 0) || (closing_ && queue_.empty() && running_ == 0);
    });
    if (stop_ || (closing_ && queue_.empty() && running_ == 0)) break;
    const int slot = free_slot();
    std::unique_ptr<Job> job(new Job);
    job->q = reqs_[queue_.front()].q;
    if (cfg_.merge) {
      // every queued request of the oldest request's pattern, in order
      std::deque<uint64_t> rest;
      for (uint64_t id : queue_) {
        if (reqs_[id].q == job->q && job->ids.size() < cfg_.max_merge)
          job->ids.push_back(id);
        else
          rest.push_back(id);
      }
      queue_.swap(rest);
    } else {
      job->ids.push_back(queue_.front());
      queue_.pop_front();
    }
    // work of the union of the ranges
    std::vector<std::pair<uint32_t, uint32_t>> rg;
    for (uint64_t id : job->ids) rg.push_back({reqs_[id].vb, reqs_[id].ve});
    std::sort(rg.begin(), rg.end());
    uint64_t work = 0;
    uint32_t cur_b = rg[0].first, cur_e = rg[0].second;
    for (size_t i = 1; i <= rg.size(); ++i) {
      if (i < rg.size() && rg[i].first <= cur_e) {
        cur_e = std::max(cur_e, rg[i].second);
        continue;
      }
      work += WorkUnits(cur_b, cur_e);
      if (i < rg.size()) {
        cur_b = rg[i].first;
        cur_e = rg[i].second;
      }
    }
    job->work = work;
    // admission by estimated intermediate bytes; a pattern not seen yet
    // gets one slot's share of the cap
    const double bpu = bytes_per_unit_[job->q];
    uint64_t reserve = cfg_.mem_cap / static_cast<uint64_t>(cfg_.slots);
    if (bpu >= 0) {
      const double est = bpu * static_cast<double>(work);
      reserve = est >= static_cast<double>(cfg_.mem_cap) ? cfg_.mem_cap : static_cast<uint64_t>(est);
    }
    reserve = std::max(reserve, cfg_.min_reserve);
    if (cfg_.slots == 1) reserve = cfg_.mem_cap;  // one pass at a time: the whole budget
    job->reserve = reserve;
    cv_dispatch_.wait(lk, [&] { return stop_ || reserved_ + reserve <= cfg_.mem_cap; });
    if (stop_) break;
    reserved_ += reserve;
    running_++;
    stats_.peak_reserved = std::max(stats_.peak_reserved, reserved_);
    stats_.peak_concurrent = std::max(stats_.peak_concurrent, running_);
    for (uint64_t id : job->ids) reqs_[id].state = 1;
    slot_busy_[slot] = true;
    slot_job_[slot] = std::move(job);
    cv_slot_[slot]->notify_one();
  }
  // wake the workers so they can exit
  stop_ = true;
  for (auto& cv : cv_slot_) cv->notify_all();
}

Status Server::RunJob(int slot, const Job& job, const std::vector<Request>& snap,
                      std::vector<uint64_t>* counts, std::string* err, uint64_t* peak) {
  Context* c = ctx_[slot].get();
  c->limit = job.reserve;
  const EnginePlan* plan = &plans_[job.q];
  counts->assign(snap.size(), 0);
  *peak = 0;
  if (snap.size() == 1) {
    const Status s = RunPlan(c, graph_.get(), plan, snap[0].vb, snap[0].ve, &(*counts)[0], nullptr,
                             nullptr, err);
    *peak = c->held_peak;
    return s;
  }
  uint64_t& job_peak = *peak;
  // maximal intervals of the union; one bucketed pass per interval with the
  // requests' end points as bucket bounds
  std::vector<size_t> order(snap.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return snap[a].vb < snap[b].vb; });
  size_t i = 0;
  while (i < order.size()) {
    uint32_t lo = snap[order[i]].vb, hi = snap[order[i]].ve;
    size_t j = i + 1;
    while (j < order.size() && snap[order[j]].vb <= hi) hi = std::max(hi, snap[order[j++]].ve);
    std::vector<uint32_t> bounds;
    for (size_t k = i; k < j; ++k) {
      bounds.push_back(snap[order[k]].vb);
      bounds.push_back(snap[order[k]].ve);
    }
    std::sort(bounds.begin(), bounds.end());
    bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
    std::vector<uint64_t> bucket(bounds.size() - 1);
    uint64_t total = 0;
    const Status s = RunPlanBuckets(c, graph_.get(), plan, bounds.data(), bucket.size(), bucket.data(),
                                    &total, nullptr, nullptr, err);
    // each interval's run restarts the context's high-water mark; the job's
    // peak is the largest of them
    job_peak = std::max(job_peak, c->held_peak);
    if (s != kOk) return s;
    // prefix over buckets: count of [bounds[a], bounds[b]) = pre[b] - pre[a]
    std::vector<uint64_t> pre(bucket.size() + 1, 0);
    for (size_t k = 0; k < bucket.size(); ++k) pre[k + 1] = pre[k] + bucket[k];
    for (size_t k = i; k < j; ++k) {
      const Request& r = snap[order[k]];
      const size_t a = std::lower_bound(bounds.begin(), bounds.end(), r.vb) - bounds.begin();
      const size_t b = std::lower_bound(bounds.begin(), bounds.end(), r.ve) - bounds.begin();
      (*counts)[order[k]] = pre[b] - pre[a];
    }
    (void)lo;
    i = j;
  }
  return kOk;
}

void Server::Work(int slot) {
  std::unique_lock<std::mutex> lk(mu_);
  while (true) {
    cv_slot_[slot]->wait(lk, [&] { return stop_ || slot_job_[slot] != nullptr; });
    if (!slot_job_[slot]) break;  // stopping with nothing assigned
    std::unique_ptr<Job> job = std::move(slot_job_[slot]);
    std::vector<Request> snap;
    for (uint64_t id : job->ids) snap.push_back(reqs_[id]);
    lk.unlock();
    std::vector<uint64_t> counts;
    std::string err;
    Status s = kOk;
    uint64_t peak = 0;
    try {
      s = RunJob(slot, *job, snap, &counts, &err, &peak);
    } catch (const Error& e) {
      s = e.status;
      err = e.msg;
    } catch (const std::bad_alloc&) {
      s = kErrOom;
      err = "host memory exhausted";
    } catch (const std::exception& e) {
      s = kErrInvalid;
      err = e.what();
    }
    lk.lock();
    const Clock::time_point now = Clock::now();
    for (size_t k = 0; k < job->ids.size(); ++k) {
      Request& r = reqs_[job->ids[k]];
      r.t_done = now;
      if (s == kOk) {
        r.count = counts[k];
        r.state = 2;
      } else {
        r.state = 3;
        r.status = s;
        r.err = err;
        stats_.failed++;
      }
    }
    if (s == kOk && job->work > 0) {
      // learn bytes per work unit from what the pass held at its peak
      const double ratio = 1.25 * static_cast<double>(peak) / static_cast<double>(job->work);
      double& b = bytes_per_unit_[job->q];
      b = b < 0 ? ratio : 0.5 * b + 0.5 * ratio;
    }
    stats_.passes++;
    stats_.requests_done += job->ids.size();
    stats_.max_group = std::max<uint64_t>(stats_.max_group, job->ids.size());
    reserved_ -= job->reserve;
    running_--;
    slot_busy_[slot] = false;
    cv_done_.notify_all();
    cv_dispatch_.notify_all();
  }
}

void Server::Close() {
  if (!open_) return;
  {
    std::lock_guard<std::mutex> lk(mu_);
    closing_ = true;
  }
  cv_dispatch_.notify_all();
  if (dispatcher_.joinable()) dispatcher_.join();
  for (auto& cv : cv_slot_) cv->notify_all();
  for (auto& t : workers_)
    if (t.joinable()) t.join();
  workers_.clear();
  graph_.reset();
  ctx_.clear();
  pool_.Release();
  open_ = false;
}

}  // namespace sme
