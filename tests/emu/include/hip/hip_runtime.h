// CPU stand-in for <hip/hip_runtime.h>, used ONLY by the test build under
// tests/emu.  It lets the engine's own kernel and host sources run on the
// host (no GPU in the development container) so that index arithmetic can be
// checked under AddressSanitizer and results compared with the oracle.
//
// Execution model: every work-group runs on the calling OS thread; each of
// its work-items is a fiber (ucontext).  Wave64 collectives (__ballot,
// __shfl_*) and __syncthreads are scheduling points: a fiber parks until
// every live lane of its wave (or block) has reached the same collective,
// then all of them resume with the exchanged values.  This reproduces the
// lock-step semantics the kernels rely on.  Device memory is host memory,
// streams and events are synchronous.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

typedef int hipError_t;
enum : int { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNoDevice = 100,
             hipErrorNotReady = 600 };
typedef struct EmuStream* hipStream_t;
typedef struct EmuEvent* hipEvent_t;
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0

hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void** p, size_t n);
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
#define hipEventDisableTiming 2
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** dev, const void* sym) {
  *dev = const_cast<void*>(sym);
  return hipSuccess;
}

namespace emu {
// Test hook: while set, hipEventQuery reports recorded events as not yet
// complete (models kernels still running on another stream).
extern bool g_hold_events;
dim3 CurThread();
extern dim3 g_block_idx, g_block_dim, g_grid_dim;
void Launch(dim3 grid, dim3 block, const std::function<void()>& body);
unsigned long long WaveExchange(int kind, unsigned long long v, unsigned arg);
void BlockBarrier();
enum { kBallot = 1, kShflXor = 2, kShflUp = 3, kShflDown = 4, kReadLane = 5 };
template <typename T>
inline T FromBits(unsigned long long b) {
  T t;
  std::memcpy(&t, &b, sizeof(T));
  return t;
}
template <typename T>
inline unsigned long long ToBits(T t) {
  unsigned long long b = 0;
  std::memcpy(&b, &t, sizeof(T));
  return b;
}
}  // namespace emu

#define threadIdx (::emu::CurThread())
#define blockIdx (::emu::g_block_idx)
#define blockDim (::emu::g_block_dim)
#define gridDim (::emu::g_grid_dim)
#define warpSize 64

inline unsigned long long __ballot(int p) { return emu::WaveExchange(emu::kBallot, p ? 1ull : 0ull, 0); }
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  (void)width;
  return emu::FromBits<T>(emu::WaveExchange(emu::kShflXor, emu::ToBits(v), static_cast<unsigned>(mask)));
}
template <typename T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
  (void)width;
  return emu::FromBits<T>(emu::WaveExchange(emu::kShflUp, emu::ToBits(v), delta));
}
template <typename T>
inline T __shfl_down(T v, unsigned delta, int width = 64) {
  (void)width;
  return emu::FromBits<T>(emu::WaveExchange(emu::kShflDown, emu::ToBits(v), delta));
}
// Wave-level builtins used by the kernels: a wave barrier is a collective
// scheduling point; readfirstlane returns lane 0's value.
#define __builtin_amdgcn_fence(order, scope) ((void)0)
inline void __builtin_amdgcn_wave_barrier() { (void)emu::WaveExchange(emu::kBallot, 0ull, 0); }
inline unsigned int __builtin_amdgcn_readfirstlane(unsigned int v) {
  return static_cast<unsigned int>(emu::WaveExchange(emu::kReadLane, v, 0));
}
inline int __builtin_amdgcn_readlane(int v, int lane) {
  return static_cast<int>(emu::WaveExchange(emu::kReadLane, static_cast<unsigned>(v), static_cast<unsigned>(lane)));
}
template <typename T>
inline T __shfl(T v, int src, int width = 64) {
  (void)width;
  return emu::FromBits<T>(emu::WaveExchange(emu::kReadLane, emu::ToBits(v), static_cast<unsigned>(src)));
}
inline int __clzll(long long x) { return x ? __builtin_clzll(static_cast<unsigned long long>(x)) : 64; }
inline unsigned int __popcll(unsigned long long x) { return static_cast<unsigned>(__builtin_popcountll(x)); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffs(unsigned int x) { return __builtin_ffs(static_cast<int>(x)); }
inline unsigned int min(unsigned int a, unsigned int b) { return a < b ? a : b; }
// HIP declares __umul24 with an int result (the low 32 bits of the product)
inline int __umul24(unsigned int a, unsigned int b) {
  return static_cast<int>((a & 0xffffffu) * (b & 0xffffffu));
}
inline unsigned long long __lanemask_lt() {
  const unsigned lane = threadIdx.x & 63u;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}
inline void __syncthreads() { emu::BlockBarrier(); }
inline unsigned int atomicOr(unsigned int* p, unsigned int v) {
  const unsigned int o = *p;
  *p = o | v;
  return o;
}
inline unsigned long long atomicOr(unsigned long long* p, unsigned long long v) {
  const unsigned long long o = *p;
  *p = o | v;
  return o;
}
inline unsigned int atomicCAS(unsigned int* p, unsigned int cmp, unsigned int v) {
  const unsigned int o = *p;
  if (o == cmp) *p = v;
  return o;
}
inline unsigned int atomicAdd(unsigned int* p, unsigned int v) {
  const unsigned int o = *p;
  *p = o + v;
  return o;
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  const unsigned long long o = *p;
  *p = o + v;
  return o;
}

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  ::emu::Launch((grid), (block), [=]() { kernel(__VA_ARGS__); })
