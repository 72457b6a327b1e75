// Per-frame neighbour encoding (prefetch_adjacency_kernel, rasterize.py:580-606):
// diff[e] = binary16(c_adj - c, r_adj), round-to-nearest-even.
#include "pf_common.hpp"

namespace pf {

__device__ __forceinline__ uint32_t f2h(float x) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}

// A wave takes 64 consecutive sites and deals their edges to its lanes in
// order (coalesced index reads and diff writes); a lane finds the site of
// its edge by binary search over the wave's 65 offsets in LDS.
__global__ void __launch_bounds__(256)
    adj_encode_kernel(const float4* __restrict__ spheres, int n, const int32_t* __restrict__ adj,
                      const int32_t* __restrict__ off, int64_t n_edges, uint2* __restrict__ diff) {
  __shared__ int s_off[4][WAVE + 1];
  __shared__ float4 s_sph[4][WAVE];
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int s0 = (blockIdx.x * 4 + wv) * WAVE;
  if (s0 >= n) return;
  const int ns = min(WAVE, n - s0);
  int* so = s_off[wv];
  float4* sp = s_sph[wv];
  if (lane < ns) {
    sp[lane] = spheres[s0 + lane];
    so[lane] = off[s0 + lane];
  }
  if (lane == 0) so[ns] = off[s0 + ns];  // ns may be 64: one past the last lane
  wave_sync();
  // offsets past the edge arrays (an inconsistent CSR) end the walk there
  const int e0 = so[0], e1 = (int)min((int64_t)so[ns], n_edges);
  for (int e = e0 + lane; e < e1; e += WAVE) {
    // largest l with so[l] <= e: the site owning edge e
    int l = 0;
#pragma unroll
    for (int step = WAVE / 2; step >= 1; step >>= 1)
      if (l + step < ns && so[l + step] <= e) l += step;
    const float4 c = sp[l];
    const float4 a = spheres[adj[e]];
    diff[e] = make_uint2(f2h(a.x - c.x) | (f2h(a.y - c.y) << 16), f2h(a.z - c.z) | (f2h(a.w) << 16));
  }
}

}  // namespace pf

using namespace pf;

extern "C" int pf_adj_encode(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                             const int32_t* adjacency_offsets, int64_t n_edges, uint16_t* diff, void* stream) {
  if (n_sites < 0 || n_edges < 0 || (n_sites > 0 && (!spheres || !adjacency_offsets)))
    return set_error(PF_ERR_ARG, "pf_adj_encode: bad arguments");
  if (n_edges > 0 && (!adjacency || !diff))
    return set_error(PF_ERR_ARG, "pf_adj_encode: %lld edges but a NULL adjacency or diff", (long long)n_edges);
  if (n_sites == 0 || n_edges == 0) return PF_OK;  // an empty adjacency is allowed: no site then reads it
  adj_encode_kernel<<<(n_sites + 4 * WAVE - 1) / (4 * WAVE), 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const float4*>(spheres), n_sites, adjacency, adjacency_offsets, n_edges,
      reinterpret_cast<uint2*>(diff));
  PF_HIP(hipGetLastError());
  return PF_OK;
}
