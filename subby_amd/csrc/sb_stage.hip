// Model-record upload: one copy of the staged records, then one scatter.
//
// sb_model_put stages each model's device record in pinned host memory; the
// records land wherever the model arena has room (fresh chunk space or a
// reused record), so a batch of new models is many small, scattered
// destinations.  Instead of one host->device copy per record, the staging
// buffer (records + a table of runs) goes up in one copy and this kernel
// moves every run to its arena address: one workgroup per run, 16-byte
// vector loads and stores when the run allows them.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sb_plan.h"

__global__ __launch_bounds__(256) void sb_scatter_runs_kernel(const char* __restrict__ stage,
                                                              const SbScatterRun* __restrict__ runs,
                                                              int nruns) {
  for (int r = blockIdx.x; r < nruns; r += gridDim.x) {
    const SbScatterRun q = runs[r];
    char* dst = reinterpret_cast<char*>(q.dst);
    const char* src = stage + q.src;
    if (((q.dst | q.src | q.len) & 15u) == 0) {
      const uint64_t n = q.len >> 4;
      for (uint64_t i = threadIdx.x; i < n; i += blockDim.x)
        reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
    } else {
      for (uint64_t i = threadIdx.x; i < q.len; i += blockDim.x) dst[i] = src[i];
    }
  }
}

// `stage` is the device copy of the staging buffer (256-byte aligned), `runs`
// the device address of its run table.
extern "C" hipError_t sb_launch_scatter_runs(const char* stage, const SbScatterRun* runs, int nruns,
                                             hipStream_t s) {
  if (nruns <= 0) return hipSuccess;
  const int grid = nruns < 2048 ? nruns : 2048;
  hipLaunchKernelGGL(sb_scatter_runs_kernel, dim3(grid), dim3(256), 0, s, stage, runs, nruns);
  return hipGetLastError();
}
