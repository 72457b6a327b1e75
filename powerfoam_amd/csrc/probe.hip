// Measurement kernel: the VALU issue rate the compositing kernels are
// compared with. Each lane runs eight independent FP32 FMA chains, so every
// wave issues one v_fma_f32 per cycle its SIMD allows and never waits on a
// result; the wave-instruction count per launch is known exactly.
#include "pf_common.hpp"

namespace pf {

constexpr int PROBE_THREADS = 256;
constexpr int PROBE_CHAINS = 8;

__global__ void __launch_bounds__(PROBE_THREADS) valu_probe_kernel(int iters, float seed, float* __restrict__ out) {
  float a[PROBE_CHAINS];
#pragma unroll
  for (int k = 0; k < PROBE_CHAINS; k++) a[k] = seed + (float)(threadIdx.x + k);
  const float m = 0.9999999f, c = 1e-7f;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < PROBE_CHAINS; k++) a[k] = __builtin_fmaf(a[k], m, c);
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < PROBE_CHAINS; k++) s += a[k];
  out[(size_t)blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

}  // namespace pf

using namespace pf;

extern "C" int64_t pf_probe_valu(int32_t n_blocks, int32_t iters, float* out, void* stream) {
  if (n_blocks < 1 || iters < 1 || !out) return set_error(PF_ERR_ARG, "pf_probe_valu: bad arguments");
  valu_probe_kernel<<<n_blocks, PROBE_THREADS, 0, as_stream(stream)>>>(iters, 1.0f, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_check(e, "valu_probe_kernel");
  // FMA wave-instructions issued by the launch
  return (int64_t)n_blocks * (PROBE_THREADS / WAVE) * PROBE_CHAINS * (int64_t)iters;
}
