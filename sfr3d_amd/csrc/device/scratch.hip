// The library's static device scratch: one file-scope array that every GPU
// the library runs on holds until the process ends.  The solve entry point
// that allocates nothing (sfr_solve_device_noalloc) keeps its scalars,
// partial sums and work vectors there instead of in context buffers.
//
// In host code the name `scratch` stands for a host-side shadow, not for the
// device memory: the device address is looked up once per device with
// hipGetSymbolAddress (the device current), kept per device, and only that
// address is handed to kernels.
#include <mutex>
#include "dev_common.hpp"

static __device__ __align__(256) unsigned char scratch[SFR_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_lock;
void *g_addr[kMaxDevices];
}  // namespace

// Device address of `scratch` on `device`, which must be the current device:
// looked up on the first call for that device, remembered after.
extern "C" int sfr_scratch_address(int device, void **ptr)
{
  if (!ptr || device < 0 || device >= kMaxDevices) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> hold(g_lock);
  if (!g_addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_addr[device] = p;
  }
  *ptr = g_addr[device];
  return hipSuccess;
}
