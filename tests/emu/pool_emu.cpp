// Test-build DevicePool: every block is its own malloc of the exact size, so
// AddressSanitizer sees the bounds of each intermediate array.  Capacity and
// accounting follow the product pool (subgraphmatchgpu_amd/csrc/pool.cpp).
#include <cstdlib>

#include "../../subgraphmatchgpu_amd/csrc/pool.h"

namespace sme {

DevicePool::~DevicePool() { Release(); }

bool DevicePool::Init(size_t bytes) {
  Release();
  std::lock_guard<std::mutex> lk(mu_);
  capacity_ = (bytes / kAlign) * kAlign;
  used_ = high_water_ = 0;
  base_ = reinterpret_cast<char*>(1);
  return true;
}

// The test build keeps one malloc per block here too (the product pool
// carves the external array; this one only enforces its capacity).
void DevicePool::InitExternal(void*, size_t bytes) { Init(bytes); }

void DevicePool::Release() {
  std::lock_guard<std::mutex> lk(mu_);
  ReclaimLocked(nullptr, true);
  for (hipEvent_t e : spare_events_) (void)hipEventDestroy(e);
  spare_events_.clear();
  for (auto& kv : live_) std::free(reinterpret_cast<void*>(kv.first));
  live_.clear();
  free_by_off_.clear();
  free_by_size_.clear();
  capacity_ = used_ = high_water_ = 0;
  base_ = nullptr;
}

void DevicePool::InsertFree(size_t, size_t) {}
void DevicePool::EraseFree(size_t, size_t) {}

size_t DevicePool::LargestFreeLocked() const { return capacity_ - used_; }

void* DevicePool::AllocLocked(size_t bytes) {
  if (bytes == 0 || !base_) return nullptr;
  const size_t len = (bytes + kAlign - 1) / kAlign * kAlign;
  if (used_ + len > capacity_) return nullptr;
  void* p = std::malloc(bytes);
  if (!p) return nullptr;
  live_[reinterpret_cast<size_t>(p)] = len;
  used_ += len;
  if (used_ > high_water_) high_water_ = used_;
  return p;
}

void DevicePool::FreeLocked(void* p) {
  if (!p) return;
  auto it = live_.find(reinterpret_cast<size_t>(p));
  if (it == live_.end()) return;
  used_ -= it->second;
  live_.erase(it);
  std::free(p);
}

}  // namespace sme
