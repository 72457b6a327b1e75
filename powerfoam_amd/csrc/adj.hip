// Per-frame neighbour encoding (prefetch_adjacency_kernel, rasterize.py:580-606):
// diff[e] = binary16(c_adj - c, r_adj), round-to-nearest-even.
#include "pf_common.hpp"

namespace pf {

__device__ __forceinline__ uint32_t f2h(float x) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}

__global__ void __launch_bounds__(256)
    adj_encode_kernel(const float4* __restrict__ spheres, int n, const int32_t* __restrict__ adj,
                      const int32_t* __restrict__ off, uint2* __restrict__ diff) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float4 c = spheres[s];
  const int e1 = off[s + 1];
  for (int e = off[s]; e < e1; e++) {
    const float4 a = spheres[adj[e]];
    diff[e] = make_uint2(f2h(a.x - c.x) | (f2h(a.y - c.y) << 16), f2h(a.z - c.z) | (f2h(a.w) << 16));
  }
}

}  // namespace pf

using namespace pf;

extern "C" int pf_adj_encode(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                             const int32_t* adjacency_offsets, uint16_t* diff, void* stream) {
  if (n_sites < 0 || (n_sites > 0 && (!spheres || !adjacency_offsets)))
    return set_error(PF_ERR_ARG, "pf_adj_encode: bad arguments");
  if (n_sites == 0) return PF_OK;
  if (!adjacency || !diff) {
    // an empty adjacency is allowed; the kernel never dereferences then
  }
  adj_encode_kernel<<<(n_sites + 255) / 256, 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const float4*>(spheres), n_sites, adjacency, adjacency_offsets, reinterpret_cast<uint2*>(diff));
  PF_HIP(hipGetLastError());
  return PF_OK;
}
