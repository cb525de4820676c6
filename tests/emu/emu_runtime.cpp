// Fiber-based CPU execution of HIP kernels for the test build (see
// include/hip/hip_runtime.h in this directory).  Test infrastructure only.
#include <hip/hip_runtime.h>
#include <setjmp.h>
#include <ucontext.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

namespace emu {

dim3 g_block_idx, g_block_dim, g_grid_dim;

namespace {

constexpr size_t kStack = 512 * 1024;

enum State { kRunnable, kWaitWave, kWaitBlock, kDone };

struct Lane {
  ucontext_t ctx;     // used once, to enter the fiber the first time
  jmp_buf jb;         // later switches avoid swapcontext's signal-mask syscalls
  bool started = false;
  char* stack = nullptr;
  State state = kDone;
  int kind = 0;
  unsigned long long val = 0;
  unsigned arg = 0;
  unsigned long long result = 0;
  unsigned tid = 0;
};

std::vector<Lane> g_lanes;
ucontext_t g_sched;
jmp_buf g_sched_jb;
int g_cur = -1;
const std::function<void()>* g_body = nullptr;
#if EMU_ASAN
void* g_fake = nullptr;
const void* g_sched_bottom = nullptr;
size_t g_sched_size = 0;
#endif

void SwitchToLane(int i) {
  g_cur = i;
  Lane& l = g_lanes[i];
#if EMU_ASAN
  __sanitizer_start_switch_fiber(&g_fake, l.stack, kStack);
#endif
  if (!_setjmp(g_sched_jb)) {
    if (!l.started) {
      l.started = true;
      swapcontext(&g_sched, &l.ctx);  // enters Entry(); it longjmps back
    } else {
      _longjmp(l.jb, 1);
    }
  }
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(g_fake, &g_sched_bottom, &g_sched_size);
#endif
  g_cur = -1;
}

void SwitchToSched() {
  Lane& l = g_lanes[g_cur];
#if EMU_ASAN
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, g_sched_bottom, g_sched_size);
#endif
  if (!_setjmp(l.jb)) _longjmp(g_sched_jb, 1);
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
}

// A fiber lives for the whole process: it runs the current kernel body once
// per work-group it is handed, then parks.
void Entry() {
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
#endif
  while (true) {
    (*g_body)();
    g_lanes[g_cur].state = kDone;
    SwitchToSched();
  }
}

}  // namespace

dim3 CurThread() { return dim3(g_lanes[g_cur].tid, 0, 0); }

unsigned long long WaveExchange(int kind, unsigned long long v, unsigned arg) {
  Lane& l = g_lanes[g_cur];
  l.state = kWaitWave;
  l.kind = kind;
  l.val = v;
  l.arg = arg;
  SwitchToSched();
  return g_lanes[g_cur].result;
}

void BlockBarrier() {
  g_lanes[g_cur].state = kWaitBlock;
  SwitchToSched();
}

void Launch(dim3 grid, dim3 block, const std::function<void()>& body) {
  const unsigned nthreads = block.x;
  if (nthreads == 0 || nthreads > 1024 || grid.x == 0) {
    std::fprintf(stderr, "emu: bad launch shape\n");
    std::abort();
  }
  if (g_lanes.size() < nthreads) {
    g_lanes.resize(nthreads);
    for (auto& l : g_lanes)
      if (!l.stack) l.stack = static_cast<char*>(std::malloc(kStack));
  }
  g_body = &body;
  g_block_dim = block;
  g_grid_dim = grid;
  const unsigned nwaves = (nthreads + 63) / 64;
  const unsigned long long nblocks = 1ull * grid.x * (grid.y ? grid.y : 1) * (grid.z ? grid.z : 1);
  for (unsigned long long b = 0; b < nblocks; ++b) {
    // blocks in x-fastest order, as the hardware numbers them
    g_block_idx = dim3(static_cast<unsigned>(b % grid.x), static_cast<unsigned>((b / grid.x) % (grid.y ? grid.y : 1)),
                       static_cast<unsigned>(b / (1ull * grid.x * (grid.y ? grid.y : 1))));
    for (unsigned t = 0; t < nthreads; ++t) {
      Lane& l = g_lanes[t];
      l.tid = t;
      l.state = kRunnable;
      if (!l.started) {
        getcontext(&l.ctx);
        l.ctx.uc_stack.ss_sp = l.stack;
        l.ctx.uc_stack.ss_size = kStack;
        l.ctx.uc_link = nullptr;
        makecontext(&l.ctx, Entry, 0);
      }
    }
    while (true) {
      bool any_live = false;
      for (unsigned t = 0; t < nthreads; ++t)
        if (g_lanes[t].state == kRunnable) SwitchToLane(static_cast<int>(t));
      // resolve wave collectives
      bool progressed = false;
      for (unsigned w = 0; w < nwaves; ++w) {
        const unsigned lo = w * 64, hi = std::min(nthreads, lo + 64);
        bool all_wait = true, any_wait = false;
        int kind = -1;
        for (unsigned t = lo; t < hi; ++t) {
          const Lane& l = g_lanes[t];
          if (l.state == kDone) continue;
          if (l.state != kWaitWave) {
            all_wait = false;
            break;
          }
          any_wait = true;
          if (kind < 0) kind = l.kind;
          if (kind != l.kind) {
            std::fprintf(stderr, "emu: lanes of one wave reached different collectives\n");
            std::abort();
          }
        }
        if (!all_wait || !any_wait) continue;
        unsigned long long vals[64] = {0};
        bool live[64] = {false};
        for (unsigned t = lo; t < hi; ++t) {
          vals[t - lo] = g_lanes[t].val;
          live[t - lo] = g_lanes[t].state == kWaitWave;
        }
        unsigned long long ballot = 0;
        for (unsigned i = 0; i < 64; ++i)
          if (live[i] && vals[i]) ballot |= 1ull << i;
        for (unsigned t = lo; t < hi; ++t) {
          Lane& l = g_lanes[t];
          if (l.state != kWaitWave) continue;
          const unsigned i = t - lo;
          if (kind == kBallot) {
            l.result = ballot;
          } else if (kind == kShflXor) {
            const unsigned src = i ^ l.arg;
            l.result = (src < 64 && src < hi - lo) ? vals[src] : vals[i];
          } else if (kind == kShflDown) {
            const unsigned src = i + l.arg;
            l.result = src < hi - lo ? vals[src] : vals[i];
          } else if (kind == kReadLane) {
            l.result = vals[l.arg < hi - lo ? l.arg : 0];
          } else {
            l.result = i >= l.arg ? vals[i - l.arg] : vals[i];
          }
          l.state = kRunnable;
        }
        progressed = true;
      }
      // resolve block barrier
      bool all_block = true, any_block = false;
      for (unsigned t = 0; t < nthreads; ++t) {
        const Lane& l = g_lanes[t];
        if (l.state == kDone) continue;
        any_live = true;
        if (l.state != kWaitBlock) all_block = false;
        else any_block = true;
      }
      if (all_block && any_block) {
        for (unsigned t = 0; t < nthreads; ++t)
          if (g_lanes[t].state == kWaitBlock) g_lanes[t].state = kRunnable;
        progressed = true;
      }
      if (!any_live) break;
      bool runnable = false;
      for (unsigned t = 0; t < nthreads; ++t) runnable = runnable || g_lanes[t].state == kRunnable;
      if (!runnable && !progressed) {
        std::fprintf(stderr, "emu: deadlock in block %u\n", b);
        std::abort();
      }
    }
  }
  g_body = nullptr;
}

}  // namespace emu

// ---------------------------------------------------------------------------
// Runtime API on host memory.
struct EmuStream {};
struct EmuEvent {
  std::chrono::steady_clock::time_point t;
};

hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emu error"; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new EmuStream;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
  *p = std::malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (n) std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, n, k);
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  if (n) std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new EmuEvent;
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
// work is synchronous here, so every recorded event has completed
namespace emu {
bool g_hold_events = false;
}
hipError_t hipEventQuery(hipEvent_t) { return emu::g_hold_events ? hipErrorNotReady : hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }

hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
