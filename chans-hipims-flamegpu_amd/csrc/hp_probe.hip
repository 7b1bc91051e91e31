// Issue-rate probe for the VALU roofline (no counterpart in the reference).
//
// The flux and single-pass kernels are bound by vector-instruction issue,
// not by HBM, so their roofline needs the rate at which this GPU retires
// ordinary (non-packed) FP32 VALU instructions.  k_probe_fma runs eight
// independent FMA chains per lane for a fixed count; timing a launch gives
// FMA lane-instructions per second, the denominator of the VALU fraction.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipims_abi.h"
#include "hp_common.hpp"

namespace hp {

constexpr int kProbeIters = 4096;  // x 8 chains = 32768 FMAs per lane

__global__ __launch_bounds__(256) void k_probe_fma(float* out, float a, float b) {
  float x0 = threadIdx.x * 1e-3f, x1 = x0 + 1.f, x2 = x0 + 2.f, x3 = x0 + 3.f;
  float x4 = x0 + 4.f, x5 = x0 + 5.f, x6 = x0 + 6.f, x7 = x0 + 7.f;
  for (int i = 0; i < kProbeIters; ++i) {
    x0 = __builtin_fmaf(x0, a, b);
    x1 = __builtin_fmaf(x1, a, b);
    x2 = __builtin_fmaf(x2, a, b);
    x3 = __builtin_fmaf(x3, a, b);
    x4 = __builtin_fmaf(x4, a, b);
    x5 = __builtin_fmaf(x5, a, b);
    x6 = __builtin_fmaf(x6, a, b);
    x7 = __builtin_fmaf(x7, a, b);
  }
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}

}  // namespace hp

using namespace hp;

extern "C" {

int64_t hp_probe_fma_per_lane(void) { return 8LL * kProbeIters; }

int hp_probe_fma(int32_t blocks, float* out, void* stream) {
  if (blocks <= 0 || !out) return fail(HP_EINVAL, "blocks must be > 0 and out non-null (blocks * 256 floats)");
  hipLaunchKernelGGL(k_probe_fma, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out, 0.999f, 1e-3f);
  return check_launch("k_probe_fma");
}

}  // extern "C"
