// pool.h -- device memory pool for the engine's intermediates.
//
// Replaces the reference's cnmem-backed CudaContext (src/gpu/utils/context/
// CudaContext.cuh:83-98, 315-338) and its byte accounting DeviceMemoryInfo
// (CudaContextUtils.h:7-56).  One hipMalloc'd arena per context, carved with
// a best-fit free list that coalesces neighbours on release.  All kernels of
// a context run on one stream, so a block released on the host may be handed
// out again immediately: any later kernel that writes it is ordered after
// every earlier kernel that read it.
//
// Exceeding the arena is an error (kErrOom), never a silent fallback to
// managed memory (the reference spills to cudaMallocManaged,
// CudaContext.cuh:152-163).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <set>
#include <utility>

namespace sme {

class DevicePool {
 public:
  DevicePool() = default;
  ~DevicePool();
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;

  // Reserves `bytes` of device memory.  Returns false if hipMalloc fails.
  bool Init(size_t bytes);
  void Release();

  // nullptr when the request does not fit; `bytes` == 0 returns nullptr.
  void* Alloc(size_t bytes);
  void Free(void* p);

  size_t Capacity() const { return capacity_; }
  size_t Used() const { return used_; }
  size_t Available() const { return capacity_ - used_; }
  size_t HighWater() const { return high_water_; }
  size_t LargestFree() const;
  void ResetHighWater() { high_water_ = used_; }

  static constexpr size_t kAlign = 256;

 private:
  void InsertFree(size_t off, size_t len);
  void EraseFree(size_t off, size_t len);

  char* base_ = nullptr;
  size_t capacity_ = 0;
  size_t used_ = 0;
  size_t high_water_ = 0;
  std::map<size_t, size_t> free_by_off_;               // offset -> length
  std::set<std::pair<size_t, size_t>> free_by_size_;   // (length, offset)
  std::map<size_t, size_t> live_;                      // offset -> length
};

}  // namespace sme
