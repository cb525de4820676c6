// pool_shared.cpp -- locking and stream-ordered reuse of DevicePool (see
// pool.h); shared by the product pool and the CPU test build's pool.
#include "pool.h"

namespace sme {

void DevicePool::SetShared(bool on) {
  std::lock_guard<std::mutex> lk(mu_);
  shared_ = on;
}

// A deferred block goes back to the free list only once its free event has
// completed, whichever stream asks: the free list is shared, so a block
// released early for its own stream could be handed to another stream while
// kernels of the freeing stream still read it.
void DevicePool::ReclaimLocked(hipStream_t stream, bool wait) {
  (void)stream;
  size_t keep = 0;
  for (size_t i = 0; i < deferred_.size(); ++i) {
    Deferred& d = deferred_[i];
    if (wait) (void)hipEventSynchronize(d.done);
    const bool done = wait || hipEventQuery(d.done) == hipSuccess;
    if (done) {
      FreeLocked(d.p);
      spare_events_.push_back(d.done);
    } else {
      deferred_[keep++] = d;
    }
  }
  deferred_.resize(keep);
}

void* DevicePool::Alloc(size_t bytes, hipStream_t stream) {
  std::lock_guard<std::mutex> lk(mu_);
  if (!shared_ || deferred_.empty()) return AllocLocked(bytes);
  // blocks whose frees every stream has passed
  ReclaimLocked(stream, false);
  void* p = AllocLocked(bytes);
  if (!p && !deferred_.empty()) {
    // wait for the other streams to pass their frees
    ReclaimLocked(stream, true);
    p = AllocLocked(bytes);
  }
  return p;
}

void DevicePool::Free(void* p, hipStream_t stream) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  if (!shared_ || !stream) {
    FreeLocked(p);
    return;
  }
  hipEvent_t e = nullptr;
  if (!spare_events_.empty()) {
    e = spare_events_.back();
    spare_events_.pop_back();
  } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
    // no event: order by waiting for the stream
    (void)hipGetLastError();
    (void)hipStreamSynchronize(stream);
    FreeLocked(p);
    return;
  }
  if (hipEventRecord(e, stream) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(stream);
    spare_events_.push_back(e);
    FreeLocked(p);
    return;
  }
  deferred_.push_back({p, e, stream});
}

size_t DevicePool::Used() const {
  std::lock_guard<std::mutex> lk(mu_);
  return used_;
}

size_t DevicePool::Available() const {
  std::lock_guard<std::mutex> lk(mu_);
  return capacity_ - used_;
}

size_t DevicePool::HighWater() const {
  std::lock_guard<std::mutex> lk(mu_);
  return high_water_;
}

size_t DevicePool::LargestFree() const {
  std::lock_guard<std::mutex> lk(mu_);
  return LargestFreeLocked();
}

void DevicePool::ResetHighWater() {
  std::lock_guard<std::mutex> lk(mu_);
  high_water_ = used_;
}

}  // namespace sme
