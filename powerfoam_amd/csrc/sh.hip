// Degree-3 spherical-harmonic texel colour, forward and backward.
//
// Fills the slot of the reference's per-texel colour pre-pass
// (SphericalVoronoi, color_fn.py:54-184): for texel t of site n the view
// direction is d = normalize(texel_site - eye) (color_fn.py:72) and
//   rgb = max(0, sh[n,0] + sum_{k=1..15} Y_k(d) sh[n,k]),
// with the real-SH constants of sh_utils.py:4-49 and the DC coefficient taken
// as the base colour. As the reference's colour head, no gradient flows to the
// texel positions (color_fn.py:181-184).
//
// One thread per site: the 48 coefficients are read once and applied to the
// site's T texels, so HBM traffic is 192 B (SH) + 12 T (texel sites) in and
// 12 T (rgb) out per site; the backward reads the same plus 12 T of incoming
// gradient and writes the 192 B coefficient gradient. The per-texel arrays
// of a workgroup's sites are one contiguous block; they pass through LDS so
// that HBM sees whole-line reads and writes instead of 12-byte pieces at a
// 12 T stride (rows padded to 3 T + 1 words: conflict-free LDS access).
#include "pf_common.hpp"
#include "sh_basis.hpp"

namespace pf {

constexpr int SH_SITES = 128;                        // sites per workgroup
constexpr int SH_ROW = 3 * PF_MAX_TEXELS + 1;        // LDS words per site (max)

// global block [s0 * 3T, (s0 + nb) * 3T) <-> LDS rows of 3T + 1 words
// Walks i = threadIdx.x, +SH_SITES, ... < nb * 3T, with the LDS index of i
// kept incrementally (one division per thread, not per element).
template <bool IN>
__device__ __forceinline__ void block_copy(float* lds, float* g, int s0, int nb, int T) {
  const int w = 3 * T, n = nb * w;
  const int dq = SH_SITES / w, dr = SH_SITES % w;
  int row = (int)threadIdx.x / w, col = (int)threadIdx.x % w;
  float* p = g + (size_t)s0 * w;
  for (int i = threadIdx.x; i < n; i += SH_SITES) {
    if (IN)
      lds[row * (w + 1) + col] = p[i];
    else
      p[i] = lds[row * (w + 1) + col];
    row += dq;
    col += dr;
    if (col >= w) {
      col -= w;
      row++;
    }
  }
}

__device__ __forceinline__ void block_in(float* lds, const float* g, int s0, int nb, int T) {
  block_copy<true>(lds, const_cast<float*>(g), s0, nb, T);
}

__device__ __forceinline__ void block_out(float* g, float* lds, int s0, int nb, int T) {
  block_copy<false>(lds, g, s0, nb, T);
}

__device__ __forceinline__ void load_sh(const float* sh, int s, float* c) {
  const float4* src = reinterpret_cast<const float4*>(sh + (size_t)s * 48);
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const float4 v = src[i];
    c[4 * i] = v.x; c[4 * i + 1] = v.y; c[4 * i + 2] = v.z; c[4 * i + 3] = v.w;
  }
}

__device__ __forceinline__ f3 lds_dir(const float* row, int t, f3 eye) {
  const float ts[3] = {row[3 * t], row[3 * t + 1], row[3 * t + 2]};
  return texel_dir(ts, eye);
}

__global__ void __launch_bounds__(SH_SITES)
    sh_fwd_kernel(const float* __restrict__ sh, const float* __restrict__ texel_sites, int n, int T, f3 eye,
                  float* __restrict__ rgb) {
  __shared__ float s_io[SH_SITES * SH_ROW];
  const int s0 = blockIdx.x * SH_SITES, nb = min(SH_SITES, n - s0);
  block_in(s_io, texel_sites, s0, nb, T);
  __syncthreads();
  if ((int)threadIdx.x < nb) {
    const int s = s0 + threadIdx.x;
    float c[48];
    load_sh(sh, s, c);
    float* row = s_io + threadIdx.x * (3 * T + 1);
    for (int t = 0; t < T; t++) {
      const f3 d = lds_dir(row, t, eye);
      float y[15];
      sh_basis(d, y);
      float r = c[0], g = c[1], b = c[2];
#pragma unroll
      for (int k = 0; k < 15; k++) {
        r += y[k] * c[3 * (k + 1)];
        g += y[k] * c[3 * (k + 1) + 1];
        b += y[k] * c[3 * (k + 1) + 2];
      }
      // texel t's position is dead once read: its slot takes the colour
      row[3 * t] = fmaxf(r, 0.0f);
      row[3 * t + 1] = fmaxf(g, 0.0f);
      row[3 * t + 2] = fmaxf(b, 0.0f);
    }
  }
  __syncthreads();
  block_out(rgb, s_io, s0, nb, T);
}

__global__ void __launch_bounds__(SH_SITES)
    sh_bwd_kernel(const float* __restrict__ sh, const float* __restrict__ texel_sites,
                  const float* __restrict__ grad_rgb, int n, int T, f3 eye, float* __restrict__ grad_sh) {
  __shared__ float s_ts[SH_SITES * SH_ROW];
  __shared__ float s_gr[SH_SITES * SH_ROW];
  const int s0 = blockIdx.x * SH_SITES, nb = min(SH_SITES, n - s0);
  block_in(s_ts, texel_sites, s0, nb, T);
  block_in(s_gr, grad_rgb, s0, nb, T);
  __syncthreads();
  if ((int)threadIdx.x >= nb) return;
  const int s = s0 + threadIdx.x;
  float c[48], g[48];
  load_sh(sh, s, c);
#pragma unroll
  for (int i = 0; i < 48; i++) g[i] = 0.0f;
  const float* row = s_ts + threadIdx.x * (3 * T + 1);
  const float* gi = s_gr + threadIdx.x * (3 * T + 1);
  for (int t = 0; t < T; t++) {
    const f3 d = lds_dir(row, t, eye);
    float y[15];
    sh_basis(d, y);
    float r = c[0], gg = c[1], b = c[2];
#pragma unroll
    for (int k = 0; k < 15; k++) {
      r += y[k] * c[3 * (k + 1)];
      gg += y[k] * c[3 * (k + 1) + 1];
      b += y[k] * c[3 * (k + 1) + 2];
    }
    const float gr = r > 0.0f ? gi[3 * t] : 0.0f;
    const float ggg = gg > 0.0f ? gi[3 * t + 1] : 0.0f;
    const float gb = b > 0.0f ? gi[3 * t + 2] : 0.0f;
    g[0] += gr; g[1] += ggg; g[2] += gb;
#pragma unroll
    for (int k = 0; k < 15; k++) {
      g[3 * (k + 1)] += y[k] * gr;
      g[3 * (k + 1) + 1] += y[k] * ggg;
      g[3 * (k + 1) + 2] += y[k] * gb;
    }
  }
  float4* dst = reinterpret_cast<float4*>(grad_sh + (size_t)s * 48);
#pragma unroll
  for (int i = 0; i < 12; i++) dst[i] = make_float4(g[4 * i], g[4 * i + 1], g[4 * i + 2], g[4 * i + 3]);
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_sh_fwd(const float* sh, const float* texel_sites, int32_t n_sites, int32_t n_texels, const float* eye,
              float* texel_rgb, void* stream) {
  if (n_sites < 0 || n_texels < 1 || n_texels > PF_MAX_TEXELS || !eye)
    return set_error(PF_ERR_ARG, "pf_sh_fwd: bad arguments");
  if (n_sites == 0) return PF_OK;
  if (!sh || !texel_sites || !texel_rgb) return set_error(PF_ERR_ARG, "pf_sh_fwd: null array");
  sh_fwd_kernel<<<(n_sites + SH_SITES - 1) / SH_SITES, SH_SITES, 0, as_stream(stream)>>>(sh, texel_sites, n_sites, n_texels,
                                                                      f3{eye[0], eye[1], eye[2]}, texel_rgb);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_sh_bwd(const float* sh, const float* texel_sites, const float* grad_texel_rgb, int32_t n_sites,
              int32_t n_texels, const float* eye, float* grad_sh, void* stream) {
  if (n_sites < 0 || n_texels < 1 || n_texels > PF_MAX_TEXELS || !eye)
    return set_error(PF_ERR_ARG, "pf_sh_bwd: bad arguments");
  if (n_sites == 0) return PF_OK;
  if (!sh || !texel_sites || !grad_texel_rgb || !grad_sh) return set_error(PF_ERR_ARG, "pf_sh_bwd: null array");
  sh_bwd_kernel<<<(n_sites + SH_SITES - 1) / SH_SITES, SH_SITES, 0, as_stream(stream)>>>(sh, texel_sites, grad_texel_rgb, n_sites,
                                                                      n_texels, f3{eye[0], eye[1], eye[2]}, grad_sh);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
