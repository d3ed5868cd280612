// VALU issue ceiling of the device, measured with a kernel of known
// instruction count: every thread runs eight independent float32 FMA chains
// (v_fma_f32, one VALU instruction each, no memory traffic inside the loop),
// 64 FMAs per loop iteration.  The rate it reaches in wave-instructions per
// second is the ceiling against which bench.py reports the executed VALU
// utilisation of the sky scan (its instruction count from rocprofv3's
// SQ_INSTS_VALU over its measured time).
#include <hip/hip_runtime.h>

#include "pycwb_amd.h"

namespace {

constexpr int kProbeBlock = 256;

__global__ __launch_bounds__(kProbeBlock) void k_valu_probe(float* out, int iters, float m,
                                                            float c) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = __builtin_fmaf(a[k], m, c);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  if (s == 12345.678f) out[blockIdx.x] = s;  // keeps the chains live; never taken
}

}  // namespace

extern "C" {

int pycwb_valu_probe(float* out, int blocks, int iters, void* stream) {
  if (!out || blocks < 1 || iters < 1) return PYCWB_ERR_ARG;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_valu_probe, dim3((unsigned)blocks), dim3(kProbeBlock), 0,
                     (hipStream_t)stream, out, iters, 0.999f, 1e-4f);
  return hipGetLastError() == hipSuccess ? PYCWB_OK : PYCWB_ERR_HIP;
}

}  // extern "C"
