// Stream-ordered reuse in the shared DevicePool: a block freed by one
// stream must stay out of the free list until the event recorded at its
// free has completed, even after the freeing stream allocates again.
//   pool_test  -> prints "ok" or the failed check, exit status 0 / 1
#include <cstdio>

#include "../../subgraphmatchgpu_amd/csrc/pool.h"

#define CHECK(c)                                         \
  do {                                                   \
    if (!(c)) {                                          \
      std::printf("failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                          \
    }                                                    \
  } while (0)

int main() {
  sme::DevicePool pool;
  CHECK(pool.Init(1 << 20));
  pool.SetShared(true);
  hipStream_t a = nullptr, b = nullptr;
  CHECK(hipStreamCreateWithFlags(&a, hipStreamNonBlocking) == hipSuccess);
  CHECK(hipStreamCreateWithFlags(&b, hipStreamNonBlocking) == hipSuccess);
  emu::g_hold_events = true;  // kernels of both streams still running
  void* p = pool.Alloc(4096, a);
  CHECK(p != nullptr);
  pool.Free(p, a);
  CHECK(pool.Used() == 4096);  // deferred, still held
  void* q = pool.Alloc(8192, a);  // the freeing stream allocates again
  CHECK(q != nullptr);
  CHECK(pool.Used() == 4096 + 8192);  // p still held: another stream could get it
  void* r = pool.Alloc(4096, b);
  CHECK(r != nullptr && r != p);
  emu::g_hold_events = false;  // stream a has passed the free
  pool.Free(q, a);
  pool.Free(r, b);
  void* s = pool.Alloc(256, b);
  CHECK(s != nullptr);
  CHECK(pool.Used() == 256);  // every completed free returned
  pool.Free(s, b);
  // a full pool waits for pending frees instead of failing
  emu::g_hold_events = true;
  void* big = pool.Alloc(1 << 20, a);
  CHECK(big != nullptr);
  pool.Free(big, a);
  void* again = pool.Alloc(1 << 20, b);
  CHECK(again != nullptr);
  emu::g_hold_events = false;
  pool.Free(again, b);
  pool.Release();
  std::printf("ok\n");
  return 0;
}
