// smc_probe.hip -- the plain FP32 FMA rate of this build on the device: a
// chain of independent v_fma_f32 per lane (packed FP32 is off for all code
// here), the figure the VALU kernels' rooflines are read against beside
// MI355X's published FP32 number.
#include <hip/hip_runtime.h>

#include "smc_internal.h"

namespace smc {

__global__ __launch_bounds__(256) void k_fma_chain(float* out, int iters, float a, float b) {
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = (float)threadIdx.x * 1e-3f + (float)k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
#pragma unroll
      for (int k = 0; k < 8; ++k) x[k] = fmaf(x[k], a, b);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += x[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// blocks x 256 lanes, each 128 x iters FMAs; out holds blocks x 256 floats
int launch_fma_chain(float* out, int blocks, int iters, void* stream) {
  if (blocks <= 0 || iters <= 0) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fma_chain, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, iters, 0.999f, 1e-4f);
  return (int)hipGetLastError();
}

}  // namespace smc
