// Geometry utilities of the training step: Morton site ordering
// (morton_code_kernel / morton_sort, geometry.py:7-85) and the
// interpenetration regulariser (interpenetration_kernel and the Warp adjoint
// of InterpenetrationFunction, geometry.py:88-189).
//
// Morton codes follow the reference's float32 arithmetic exactly: the box of
// the points (min / max per axis), extents floored at 1e-6, (p - lower) /
// extent with IEEE division, times 2^21, clamped to [0, 2^21 - 1], truncated,
// bits interleaved x-major (x at bit 3k+2, y at 3k+1, z at 3k). The ordering
// is a stable LSD radix sort of (code, index) pairs (rocPRIM), as Warp's
// radix_sort_pairs with the identity as values: ties keep index order.
//
// The regulariser: area_i = sum over i's neighbours j with d < r_i + r_j of
// (r_i + r_j - d)^2, d = |c_i - c_j| in float32. One wave per site, lanes over
// its edges, a wave sum; the adjoint sends g_i 2 (r_i + r_j - d) to r_i and
// r_j and -/+ g_i 2 (r_i + r_j - d) (c_i - c_j) / d to c_i / c_j (nothing
// through the direction when d = 0), reducing i's share over the wave
// before one atomic per component; entries whose reverse is in the other
// list are summed by each end from its own list (no atomics).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "pf_common.hpp"

namespace pf {

__device__ __forceinline__ int geo_ord(float f) {
  const int i = __builtin_bit_cast(int, f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float geo_float(int i) { return __builtin_bit_cast(float, i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void geo_bounds_init_kernel(int* __restrict__ acc) {
  if (threadIdx.x < 3) acc[threadIdx.x] = 0x7fffffff;
  else if (threadIdx.x < 6) acc[threadIdx.x] = (int)0x80000000;
}

// min / max of the points per axis (acc: min xyz, max xyz as ordered ints).
__global__ void __launch_bounds__(256) geo_bounds_kernel(const float* __restrict__ points, int n,
                                                        int* __restrict__ acc) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int k = 0; k < 3; k++) {
      const float v = points[3 * (int64_t)i + k];
      mn[k] = fminf(mn[k], v);
      mx[k] = fmaxf(mx[k], v);
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], m, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], m, 64));
    }
  // one atomic per value per workgroup: the waves meet in LDS first
  __shared__ float s_b[4][6];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; k++) {
      s_b[w][k] = mn[k];
      s_b[w][3 + k] = mx[k];
    }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    const float v = fminf(fminf(s_b[0][k], s_b[1][k]), fminf(s_b[2][k], s_b[3][k]));
    atomicMin(&acc[k], geo_ord(v));
  } else if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    const float v = fmaxf(fmaxf(s_b[0][k], s_b[1][k]), fmaxf(s_b[2][k], s_b[3][k]));
    atomicMax(&acc[k], geo_ord(v));
  }
}

// expand_bits, geometry.py:7-15: the low 21 bits of v to every third bit.
__device__ __forceinline__ uint64_t geo_expand(uint64_t v) {
  v &= 0x00000000001FFFFFull;
  v = (v | (v << 32)) & 0x001F00000000FFFFull;
  v = (v | (v << 16)) & 0x001F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__device__ __forceinline__ uint64_t geo_quant(float p, float lower, float extent) {
  const float res = 2097152.0f;
  const float x = fminf(fmaxf(((p - lower) / extent) * res, 0.0f), res - 1.0f);
  return (uint64_t)x;
}

__global__ void __launch_bounds__(256) geo_morton_kernel(const float* __restrict__ points, int n,
                                                        const int* __restrict__ acc, uint64_t* __restrict__ codes,
                                                        int32_t* __restrict__ index) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float lo[3], ext[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = geo_float(acc[k]);
    ext[k] = fmaxf(geo_float(acc[3 + k]) - lo[k], 1e-6f);
  }
  const float* p = points + 3 * (int64_t)i;
  const uint64_t xx = geo_expand(geo_quant(p[0], lo[0], ext[0]));
  const uint64_t yy = geo_expand(geo_quant(p[1], lo[1], ext[1]));
  const uint64_t zz = geo_expand(geo_quant(p[2], lo[2], ext[2]));
  codes[i] = xx * 4ull + yy * 2ull + zz;
  index[i] = i;
}

// ---- interpenetration ------------------------------------------------------

__device__ __forceinline__ float geo_len(float dx, float dy, float dz) { return sqrtf(dx * dx + dy * dy + dz * dz); }

__global__ void __launch_bounds__(256) geo_interp_fwd_kernel(const float4* __restrict__ spheres, int n,
                                                            const int32_t* __restrict__ adj,
                                                            const int32_t* __restrict__ off, int64_t n_edges,
                                                            float* __restrict__ areas) {
  const int i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  if (i >= n) return;
  const float4 si = spheres[i];
  const int64_t e0 = min((int64_t)off[i], n_edges), e1 = min((int64_t)off[i + 1], n_edges);
  float acc = 0.0f;
  for (int64_t e = e0 + lane; e < e1; e += WAVE) {
    const int j = adj[e];
    if ((unsigned)j >= (unsigned)n) continue;
    const float4 sj = spheres[j];
    const float d = geo_len(si.x - sj.x, si.y - sj.y, si.z - sj.z);
    const float s = si.w + sj.w;
    if (d < s) acc += (s - d) * (s - d);
  }
  acc = wave_sum(acc);
  if (lane == 0) areas[i] = acc;
}

// The adjoint. Every entry (i, j) carries two terms: area_i's derivative at
// c_i / r_i, and area_i's derivative at c_j / r_j. PAIRING 0 sends the second
// to j with atomics. When j's list holds the reverse entry (j, i) -- always
// with PAIRING 2 (symmetric lists), found by binary search with PAIRING 1
// (ascending lists without repeats) -- i instead gathers area_j's term at
// c_i / r_i through g_j (the same d and r_i + r_j), and j does the same, so
// those entries need no atomics at all.
template <int PAIRING>
__global__ void __launch_bounds__(256) geo_interp_bwd_kernel(const float4* __restrict__ spheres, int n,
                                                            const int32_t* __restrict__ adj,
                                                            const int32_t* __restrict__ off, int64_t n_edges,
                                                            const float* __restrict__ grad_areas,
                                                            float* __restrict__ grad_spheres) {
  const int i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  if (i >= n) return;
  const float g = grad_areas[i];
  const int64_t e0 = min((int64_t)off[i], n_edges), e1 = min((int64_t)off[i + 1], n_edges);
  if (e1 <= e0 || (PAIRING == 0 && g == 0.0f)) return;
  const float4 si = spheres[i];
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, gr = 0.0f;
  for (int64_t e = e0 + lane; e < e1; e += WAVE) {
    const int j = adj[e];
    if ((unsigned)j >= (unsigned)n) continue;
    const float4 sj = spheres[j];
    const float dx = si.x - sj.x, dy = si.y - sj.y, dz = si.z - sj.z;
    const float d = geo_len(dx, dy, dz);
    const float s = si.w + sj.w;
    if (!(d < s)) continue;
    bool paired = PAIRING == 2;
    if (PAIRING == 1) {
      int64_t lo = min((int64_t)off[j], n_edges), hi = min((int64_t)off[j + 1], n_edges);
      const int64_t end = hi;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (adj[mid] < i) lo = mid + 1;
        else hi = mid;
      }
      paired = lo < end && adj[lo] == i;
    }
    const float gi = g * 2.0f * (s - d);  // d area_i / d s, = -d area_i / d d
    if (paired) {
      const float c = (g + grad_areas[j]) * 2.0f * (s - d);
      gr += c;
      if (d > 0.0f) {
        const float k = c / d;
        gx -= k * dx;
        gy -= k * dy;
        gz -= k * dz;
      }
    } else if (gi != 0.0f) {
      gr += gi;
      float* gj = grad_spheres + 4 * (int64_t)j;
      atomicAdd(gj + 3, gi);
      if (d > 0.0f) {
        const float k = gi / d;
        gx -= k * dx;
        gy -= k * dy;
        gz -= k * dz;
        atomicAdd(gj + 0, k * dx);
        atomicAdd(gj + 1, k * dy);
        atomicAdd(gj + 2, k * dz);
      }
    }
  }
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  gr = wave_sum(gr);
  if (lane == 0 && (gx != 0.0f || gy != 0.0f || gz != 0.0f || gr != 0.0f)) {
    float* gi = grad_spheres + 4 * (int64_t)i;
    atomicAdd(gi + 0, gx);
    atomicAdd(gi + 1, gy);
    atomicAdd(gi + 2, gz);
    atomicAdd(gi + 3, gr);
  }
}

static size_t geo_align(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t geo_sort_temp(int32_t n) {
  size_t bytes = 0;
  const hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const uint64_t*>(nullptr),
                                                 static_cast<uint64_t*>(nullptr), static_cast<const int32_t*>(nullptr),
                                                 static_cast<int32_t*>(nullptr), (unsigned)n, 0, 63, (hipStream_t)0);
  return e == hipSuccess ? bytes : 0;
}

}  // namespace pf

using namespace pf;

extern "C" {

size_t pf_morton_workspace_size(int32_t n_points) {
  if (n_points < 0) return 0;
  const size_t n = (size_t)n_points;
  return geo_align(64) + 2 * geo_align(8 * n) + geo_align(4 * n) + geo_align(geo_sort_temp(n_points));
}

int pf_morton_sort(const float* points, int32_t n_points, void* workspace, size_t workspace_bytes, int64_t* codes,
                   int32_t* order, void* stream) {
  if (n_points < 0 || (n_points > 0 && (!points || !workspace || !order)))
    return set_error(PF_ERR_ARG, "pf_morton_sort: bad arguments");
  if (workspace_bytes < pf_morton_workspace_size(n_points))
    return set_error(PF_ERR_OVERFLOW, "pf_morton_sort: workspace below pf_morton_workspace_size");
  if (reinterpret_cast<uintptr_t>(workspace) % 256)
    return set_error(PF_ERR_ARG, "pf_morton_sort: workspace must be 256-byte aligned");
  if (n_points == 0) return PF_OK;
  hipStream_t st = as_stream(stream);
  const size_t n = (size_t)n_points;
  char* p = static_cast<char*>(workspace);
  int* acc = reinterpret_cast<int*>(p);
  p += geo_align(64);
  uint64_t* keys = reinterpret_cast<uint64_t*>(p);
  p += geo_align(8 * n);
  uint64_t* keys_sorted = reinterpret_cast<uint64_t*>(p);
  p += geo_align(8 * n);
  int32_t* index = reinterpret_cast<int32_t*>(p);
  p += geo_align(4 * n);
  size_t temp_bytes = geo_sort_temp(n_points);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  geo_bounds_init_kernel<<<1, 64, 0, st>>>(acc);
  geo_bounds_kernel<<<min(blocks, 512u), 256, 0, st>>>(points, n_points, acc);
  geo_morton_kernel<<<blocks, 256, 0, st>>>(points, n_points, acc, keys, index);
  PF_HIP(hipGetLastError());
  if (codes) PF_HIP(hipMemcpyAsync(codes, keys, 8 * n, hipMemcpyDeviceToDevice, st));
  PF_HIP(rocprim::radix_sort_pairs(p, temp_bytes, static_cast<const uint64_t*>(keys), keys_sorted,
                                   static_cast<const int32_t*>(index), order, (unsigned)n, 0, 63, st));
  return PF_OK;
}

int pf_interpenetration_fwd(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                            const int32_t* adjacency_offsets, int64_t n_edges, float* areas, void* stream) {
  if (n_sites < 0 || n_edges < 0 || (n_sites > 0 && (!spheres || !adjacency_offsets || !areas)) ||
      (n_edges > 0 && !adjacency))
    return set_error(PF_ERR_ARG, "pf_interpenetration_fwd: bad arguments");
  if (n_sites == 0) return PF_OK;
  geo_interp_fwd_kernel<<<(n_sites + 3) / 4, 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const float4*>(spheres), n_sites, adjacency, adjacency_offsets, n_edges, areas);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_interpenetration_bwd(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                            const int32_t* adjacency_offsets, int64_t n_edges, int32_t pairing,
                            const float* grad_areas, float* grad_spheres, void* stream) {
  if (pairing < 0 || pairing > 2) return set_error(PF_ERR_ARG, "pf_interpenetration_bwd: pairing must be 0, 1 or 2");
  if (n_sites < 0 || n_edges < 0 ||
      (n_sites > 0 && (!spheres || !adjacency_offsets || !grad_areas || !grad_spheres)) ||
      (n_edges > 0 && !adjacency))
    return set_error(PF_ERR_ARG, "pf_interpenetration_bwd: bad arguments");
  if (n_sites == 0) return PF_OK;
  const float4* sp = reinterpret_cast<const float4*>(spheres);
  const unsigned blocks = (unsigned)((n_sites + 3) / 4);
  hipStream_t st = as_stream(stream);
  if (pairing == 2)
    geo_interp_bwd_kernel<2><<<blocks, 256, 0, st>>>(sp, n_sites, adjacency, adjacency_offsets, n_edges, grad_areas,
                                                     grad_spheres);
  else if (pairing == 1)
    geo_interp_bwd_kernel<1><<<blocks, 256, 0, st>>>(sp, n_sites, adjacency, adjacency_offsets, n_edges, grad_areas,
                                                     grad_spheres);
  else
    geo_interp_bwd_kernel<0><<<blocks, 256, 0, st>>>(sp, n_sites, adjacency, adjacency_offsets, n_edges, grad_areas,
                                                     grad_spheres);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
