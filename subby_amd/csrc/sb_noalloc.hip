// Device memory of the no-allocation Counts entry point (sb_counts_noalloc).
//
// One file-scope array per GPU, reserved when the library's code object is
// loaded on that GPU and held until the process ends.  In host code the
// array's name is only a host-side stand-in: its device address on the
// current GPU comes from hipGetSymbolAddress (sb_create calls
// sb_internal_noalloc_scratch once per device, with that device current),
// and only that address is ever handed to copies or kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "sb_plan.h"

extern "C" int sb_internal_error(int code, const char* msg);

static __device__ __align__(256) char scratch[SB_NOALLOC_BYTES];

namespace {
std::mutex g_mu;
void* g_ptr[SB_NOALLOC_MAX_DEVICES];
}  // namespace

// Device address of this GPU's scratch array (the caller made `device`
// current).  Looked up once per device; returns SB_OK (0) or an error code.
extern "C" int sb_internal_noalloc_scratch(int device, void** ptr) {
  if (device < 0 || device >= SB_NOALLOC_MAX_DEVICES || !ptr)
    return sb_internal_error(-1, "no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char buf[200];
      std::snprintf(buf, sizeof buf, "hipGetSymbolAddress of the no-allocation scratch failed: %s",
                    hipGetErrorString(e));
      (void)hipGetLastError();
      return sb_internal_error(-5, buf);
    }
    g_ptr[device] = p;
  }
  *ptr = g_ptr[device];
  return 0;
}
