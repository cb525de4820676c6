// static_scratch.hip -- the device scratch of the no-allocation entry point
// (sme_count_static).
//
// Every intermediate of a run (frames, candidate sets, scan and batch-split
// temporaries, counters) normally comes from the context's pool, whose
// arena sme_open reserves with hipMalloc.  Callers that cannot allocate
// while they count (code under stream capture, tight serving loops) use
// this file's array instead: a file-scope device array of kScratchBytes,
// which each device the library runs on holds until the process ends.  Its
// device address is looked up once per device at sme_open, with the device
// current, and only that address is used afterwards (in host code the
// array's name is not a device address).  All calls on one device share
// the array: calls that overlap on one device race, and each call starts
// by clearing what it needs (the engine's counters are cleared by a fill
// kernel at the start of every run).
#include <mutex>
#include <string>

#include "engine.h"

namespace sme {

namespace {

constexpr uint64_t kScratchBytes = 64ull << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
  bool tried = false;
  void* ptr = nullptr;  // device address of `scratch` on this device
  std::string err;
  DevicePool pool;      // carves ptr[0, kScratchBytes)
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

}  // namespace

uint64_t StaticScratchBytes() { return kScratchBytes; }

Status InitStaticScratch(int device, std::string* err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = "device id beyond the static scratch table";
    return kErrInvalid;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceScratch& d = g_dev[device];
  if (d.tried) {
    *err = d.err;
    return d.ptr ? kOk : kErrHip;
  }
  d.tried = true;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    d.err = std::string("hipGetSymbolAddress of the static scratch failed: ") + hipGetErrorString(e);
    *err = d.err;
    return kErrHip;
  }
  d.ptr = ptr;
  d.pool.InitExternal(ptr, kScratchBytes);
  return kOk;
}

Status RunPlanNoAlloc(Context* ctx, const DevGraph* g, const EnginePlan* plan, uint32_t v_begin,
                      uint32_t v_end, uint64_t* count, uint64_t* level_counts, RunStats* stats,
                      std::string* err) {
  if (!ctx) {
    *err = "null argument";
    return kErrInvalid;
  }
  const int dev = ctx->device();
  DeviceScratch* d = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev >= 0 && dev < kMaxDevices && g_dev[dev].ptr) d = &g_dev[dev];
    else if (dev >= 0 && dev < kMaxDevices) *err = g_dev[dev].err;
  }
  if (!d) {
    if (err->empty()) *err = "the static scratch of this device is not set up";
    return kErrHip;
  }
  // the run sees the whole array as its budget; no HIP events are created
  DevicePool* own = ctx->SwapPool(&d->pool);
  const bool prof = ctx->profiling;
  const uint64_t limit = ctx->limit, held = ctx->held;
  ctx->profiling = false;
  ctx->limit = 0;
  ctx->held = 0;
  const Status s = RunPlan(ctx, g, plan, v_begin, v_end, count, level_counts, stats, err);
  ctx->SwapPool(own);
  ctx->profiling = prof;
  ctx->limit = limit;
  ctx->held = held;
  if (s == kErrOom)
    *err = "the no-allocation scratch holds " + std::to_string(kScratchBytes >> 20) +
           " MiB per device, too little for one batch of this call (" + *err + ")";
  return s;
}

}  // namespace sme
