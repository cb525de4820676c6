// pool.cpp -- see pool.h.
#include "pool.h"

#include <hip/hip_runtime.h>

namespace sme {

DevicePool::~DevicePool() { Release(); }

bool DevicePool::Init(size_t bytes) {
  Release();
  std::lock_guard<std::mutex> lk(mu_);
  bytes = (bytes / kAlign) * kAlign;
  if (bytes == 0) return true;
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  base_ = static_cast<char*>(p);
  capacity_ = bytes;
  used_ = 0;
  high_water_ = 0;
  InsertFree(0, bytes);
  return true;
}

void DevicePool::InitExternal(void* base, size_t bytes) {
  Release();
  std::lock_guard<std::mutex> lk(mu_);
  bytes = (bytes / kAlign) * kAlign;
  if (!base || bytes == 0) return;
  base_ = static_cast<char*>(base);
  external_ = true;
  capacity_ = bytes;
  used_ = 0;
  high_water_ = 0;
  InsertFree(0, bytes);
}

void DevicePool::Release() {
  std::lock_guard<std::mutex> lk(mu_);
  ReclaimLocked(nullptr, true);
  for (hipEvent_t e : spare_events_) (void)hipEventDestroy(e);
  spare_events_.clear();
  if (base_ && !external_) (void)hipFree(base_);
  base_ = nullptr;
  external_ = false;
  capacity_ = used_ = high_water_ = 0;
  free_by_off_.clear();
  free_by_size_.clear();
  live_.clear();
}

void DevicePool::InsertFree(size_t off, size_t len) {
  free_by_off_[off] = len;
  free_by_size_.insert({len, off});
}

void DevicePool::EraseFree(size_t off, size_t len) {
  free_by_off_.erase(off);
  free_by_size_.erase({len, off});
}

size_t DevicePool::LargestFreeLocked() const {
  return free_by_size_.empty() ? 0 : free_by_size_.rbegin()->first;
}

void* DevicePool::AllocLocked(size_t bytes) {
  if (bytes == 0 || !base_) return nullptr;
  const size_t len = (bytes + kAlign - 1) / kAlign * kAlign;
  auto it = free_by_size_.lower_bound({len, 0});  // best fit
  if (it == free_by_size_.end()) return nullptr;
  const size_t blk_len = it->first, blk_off = it->second;
  EraseFree(blk_off, blk_len);
  if (blk_len > len) InsertFree(blk_off + len, blk_len - len);
  live_[blk_off] = len;
  used_ += len;
  if (used_ > high_water_) high_water_ = used_;
  return base_ + blk_off;
}

void DevicePool::FreeLocked(void* p) {
  if (!p) return;
  size_t off = static_cast<size_t>(static_cast<char*>(p) - base_);
  auto lv = live_.find(off);
  if (lv == live_.end()) return;  // not ours: ignore
  size_t len = lv->second;
  live_.erase(lv);
  used_ -= len;
  // coalesce with the following free block
  auto next = free_by_off_.lower_bound(off);
  if (next != free_by_off_.end() && next->first == off + len) {
    const size_t nl = next->second;
    EraseFree(next->first, nl);
    len += nl;
  }
  // and with the preceding one
  auto prev = free_by_off_.lower_bound(off);
  if (prev != free_by_off_.begin()) {
    --prev;
    if (prev->first + prev->second == off) {
      const size_t po = prev->first, pl = prev->second;
      EraseFree(po, pl);
      off = po;
      len += pl;
    }
  }
  InsertFree(off, len);
}

}  // namespace sme
