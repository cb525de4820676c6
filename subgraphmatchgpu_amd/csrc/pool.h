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
//
// Shared mode (several contexts, each with its own stream, carving one
// arena, as the slots of a server do): a freed block returns to the free
// list only after an event recorded on the freeing stream at the free has
// completed, so a kernel of the next owner (on any stream) never overwrites
// data a kernel of the old owner may still read.  All public members lock
// the pool.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <set>
#include <utility>
#include <vector>

namespace sme {

class DevicePool {
 public:
  DevicePool() = default;
  ~DevicePool();
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;

  // Reserves `bytes` of device memory.  Returns false if hipMalloc fails.
  bool Init(size_t bytes);
  // Carves `bytes` of device memory the caller owns (no hipMalloc, and
  // Release does not free it).
  void InitExternal(void* base, size_t bytes);
  void Release();

  // nullptr when the request does not fit; `bytes` == 0 returns nullptr.
  // `stream` is the stream whose kernels use the block (shared mode).
  void* Alloc(size_t bytes, hipStream_t stream = nullptr);
  void Free(void* p, hipStream_t stream = nullptr);
  void SetShared(bool on);

  size_t Capacity() const { return capacity_; }
  size_t Used() const;
  size_t Available() const;
  size_t HighWater() const;
  size_t LargestFree() const;
  void ResetHighWater();

  static constexpr size_t kAlign = 256;
  static size_t Rounded(size_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

 private:
  void* AllocLocked(size_t bytes);
  void FreeLocked(void* p);
  size_t LargestFreeLocked() const;
  // Returns deferred blocks whose free event completed; with `wait`, waits
  // for all of them.
  void ReclaimLocked(hipStream_t stream, bool wait);
  void InsertFree(size_t off, size_t len);
  void EraseFree(size_t off, size_t len);

  struct Deferred {
    void* p;
    hipEvent_t done;
    hipStream_t stream;
  };
  mutable std::mutex mu_;
  bool shared_ = false;
  std::vector<Deferred> deferred_;
  std::vector<hipEvent_t> spare_events_;

  char* base_ = nullptr;
  bool external_ = false;
  size_t capacity_ = 0;
  size_t used_ = 0;
  size_t high_water_ = 0;
  std::map<size_t, size_t> free_by_off_;               // offset -> length
  std::set<std::pair<size_t, size_t>> free_by_size_;   // (length, offset)
  std::map<size_t, size_t> live_;                      // offset -> length
};

}  // namespace sme
