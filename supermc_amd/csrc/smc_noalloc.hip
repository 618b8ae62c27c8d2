// smc_noalloc.hip -- the device scratch of smc_run_noalloc: the moment
// kernels' per-event partial sums and hot-spot rectangles, in one
// file-scope array instead of buffers the library allocates.
//
// Each GPU the library runs on holds its own copy of `scratch` until the
// process ends.  Its device address is looked up once per device (with that
// device current) and kept; host code passes that address, never the
// array's name, to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "smc_internal.h"

namespace smc {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

// Device address of this device's scratch array (nullptr and the HIP error
// in *rc on failure).
void* noalloc_scratch(int* rc) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *rc = e != hipSuccess ? (int)e : (int)hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *rc = e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
      return nullptr;
    }
    g_ptr[dev] = p;
  }
  *rc = 0;
  return g_ptr[dev];
}

}  // namespace smc
