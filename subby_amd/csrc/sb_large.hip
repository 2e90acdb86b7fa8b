// Large-alphabet path (padded alphabet 16..64): placeholder until implemented.
#include <hip/hip_runtime.h>
#include "sb_plan.h"

extern "C" long long sb_large_wave_stride(int, int, int) { return 256; }

extern "C" hipError_t sb_launch_large(int, int, bool, bool, int, int, const SbLaunch*, hipStream_t) {
  return hipErrorNotSupported;
}
