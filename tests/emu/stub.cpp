// TEST INFRASTRUCTURE -- host stand-ins for the HIP runtime calls the
// library makes, for the CPU emulation build (see emu.hpp), and the fiber
// scheduler that runs emulated workgroups.  "Device" memory is host memory
// filled with 0x7f bytes (garbage, like fresh device memory, so a read of
// memory never written shows up); copies are memcpy; a device with 4 CUs.
#include <hip/hip_runtime_api.h>
#include <ucontext.h>
#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#else
static inline void __sanitizer_start_switch_fiber(void **, const void *, size_t) {}
static inline void __sanitizer_finish_switch_fiber(void *, const void **, size_t *) {}
#endif
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <thread>
#include <vector>
#include "emu.hpp"
#include "sfr_internal.h"

namespace emu {
thread_local Idx tIdx, bIdx;
Idx gDim, bDim;
thread_local int lane;

// work-items of one workgroup run as fibers on this thread; a barrier
// (or each half of a shuffle) switches back to the scheduler
enum { MAXT = 256, STACK = 256 * 1024 };
static ucontext_t sched_ctx, item_ctx[MAXT];
static char *stacks[MAXT];
static int cur, finished[MAXT], nitems;
static const std::function<void()> *job;
static double buf[MAXT];

static void yield_to_sched()
{
  void *fs = nullptr;
  __sanitizer_start_switch_fiber(&fs, nullptr, 0);   // back to the scheduler's stack
  swapcontext(&item_ctx[cur], &sched_ctx);
  __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
}

static void item_main()
{
  __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
  (*job)();
  finished[cur] = 1;
  __sanitizer_start_switch_fiber(nullptr, nullptr, 0);
  setcontext(&sched_ctx);
}

void sync_block() { yield_to_sched(); }

double shfl(double v, int src)
{
  buf[cur] = v;
  yield_to_sched();
  const int wbase = cur & ~63;
  const double r = buf[wbase + src];
  yield_to_sched();
  return r;
}

static void run_block(unsigned b, unsigned block)
{
  nitems = (int)block;
  for (int t = 0; t < nitems; t++) {
    if (!stacks[t]) stacks[t] = (char *)malloc(STACK);
    getcontext(&item_ctx[t]);
    item_ctx[t].uc_stack.ss_sp = stacks[t];
    item_ctx[t].uc_stack.ss_size = STACK;
    item_ctx[t].uc_link = &sched_ctx;
    makecontext(&item_ctx[t], item_main, 0);
    finished[t] = 0;
  }
  bIdx = {b, 0, 0};
  for (;;) {
    int live = 0;
    for (int t = 0; t < nitems; t++) {
      if (finished[t]) continue;
      cur = t;
      tIdx = {(unsigned)t, 0, 0};
      lane = t & 63;
      void *fs = nullptr;
      __sanitizer_start_switch_fiber(&fs, stacks[t], STACK);
      swapcontext(&sched_ctx, &item_ctx[t]);
      __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
      if (!finished[t]) live++;
    }
    if (!live) break;
  }
}

static void run_grid(const std::function<void()> &f, unsigned grid, unsigned block)
{
  if (block > MAXT || block % 64) { fprintf(stderr, "emu: bad block %u\n", block); abort(); }
  gDim = {grid, 1, 1};
  bDim = {block, 1, 1};
  job = &f;
  for (unsigned b = 0; b < grid; b++) run_block(b, block);
}

static bool capturing = false;
static std::vector<std::function<void()>> *cap = nullptr;

void launch_fn(std::function<void()> f, unsigned grid, unsigned block)
{
  if (grid == 0) { fprintf(stderr, "emu: empty grid\n"); abort(); }
  if (capturing) {
    cap->push_back([f, grid, block]() { run_grid(f, grid, block); });
    return;
  }
  run_grid(f, grid, block);
}

unsigned long long clock()
{
  static unsigned long long c = 1000;
  return c += 7;
}
}  // namespace emu

struct ihipStream_t { int dummy; };
struct ihipEvent_t { double t; };
struct ihipGraph { std::vector<std::function<void()>> ops; };
struct hipGraphExec { std::vector<std::function<void()>> ops; };

static double now_ms()
{
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device allocations, so that copies can be checked the way the runtime
// checks them: a device-side range must lie inside one allocation.
static std::map<uintptr_t, size_t> g_dev;

static bool dev_range(const void *p, size_t n)
{
  const uintptr_t a = (uintptr_t)p;
  auto it = g_dev.upper_bound(a);
  if (it == g_dev.begin()) return false;
  --it;
  return a >= it->first && a + n <= it->first + it->second;
}

static hipError_t check_copy(void *d, const void *s, size_t n, hipMemcpyKind k)
{
  const bool dd = k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice;
  const bool sd = k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice;
  if (!d || !s || (dd && !dev_range(d, n)) || (sd && !dev_range(s, n))) {
    fprintf(stderr, "emu: invalid copy d=%p s=%p n=%zu kind=%d\n", d, s, n, (int)k);
    return hipErrorInvalidValue;
  }
  return hipSuccess;
}

extern "C" {
hipError_t hipMalloc(void **p, size_t n)
{
  *p = malloc(n ? n : 1);
  memset(*p, 0x7f, n);
  g_dev[(uintptr_t)*p] = n ? n : 1;
  return hipSuccess;
}
hipError_t hipFree(void *p)
{
  if (p) g_dev.erase((uintptr_t)p);
  free(p);
  return hipSuccess;
}
// A __device__ variable is host memory here; the library's one such
// variable is its static scratch, registered as a device range of its size.
hipError_t hipGetSymbolAddress(void **p, const void *sym)
{
  *p = (void *)sym;
  g_dev[(uintptr_t)sym] = SFR_SCRATCH_BYTES;
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned int)
{
  *p = malloc(n ? n : 1);
  return hipSuccess;
}
hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k)
{
  const hipError_t e = check_copy(d, s, n, k);
  if (e != hipSuccess) return e;
  memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t)
{
  if (emu::capturing) { fprintf(stderr, "emu: memcpy in capture\n"); abort(); }
  const hipError_t e = check_copy(d, s, n, k);
  if (e != hipSuccess) return e;
  memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t)
{
  if (emu::capturing) { fprintf(stderr, "emu: memset in capture\n"); abort(); }
  memset(d, v, n);
  return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
  *v = a == hipDeviceAttributeMultiprocessorCount ? 4 : 100000;
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int) { *s = new ihipStream_t; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new ihipEvent_t{0}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = now_ms(); return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(b->t - a->t); return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
  emu::capturing = true;
  emu::cap = new std::vector<std::function<void()>>();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
  emu::capturing = false;
  *g = new ihipGraph{*emu::cap};
  delete emu::cap;
  emu::cap = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
  *e = new hipGraphExec{g->ops};
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
  for (auto &f : e->ops) f();
  return hipSuccess;
}
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "hipSuccess" : "hipError"; }
hipError_t hipGetLastError(void) { return hipSuccess; }
}
