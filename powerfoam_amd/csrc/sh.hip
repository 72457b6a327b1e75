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
// gradient and writes the 192 B coefficient gradient. The per-texel arrays of
// a workgroup's sites are one contiguous block; they pass through LDS so that
// HBM sees whole-line reads and writes instead of 12-byte pieces at a 12 T
// stride (LDS rows padded to an odd word count: conflict-free). The
// coefficients are read and written as 16-byte pieces directly (staging them
// too measured slower: the LDS it takes halves the resident waves).
#include "pf_common.hpp"
#include "sh_basis.hpp"

#include <initializer_list>

namespace pf {

constexpr int SH_SITES = 128;                  // sites per workgroup
constexpr int SH_ROW = 3 * PF_MAX_TEXELS + 1;  // LDS words per site of a per-texel array (max)

// global block [s0 * w, (s0 + nb) * w) <-> LDS rows of w + 1 words (odd:
// conflict-free when each thread reads its own row). Walks i = threadIdx.x,
// +SH_SITES, ... < nb * w with the LDS index of i kept incrementally (one
// division per thread, not per element).
template <bool IN>
__device__ __forceinline__ void block_copy(float* lds, float* g, int s0, int nb, int w, bool vec) {
  if (vec && (w & 3) == 0) {
    // 16-byte pieces (a piece never straddles two rows when 4 divides w)
    const int w4 = w >> 2, n4 = nb * w4;
    const int dq = SH_SITES / w4, dr = SH_SITES % w4;
    int row = (int)threadIdx.x / w4, col = (int)threadIdx.x % w4;
    float4* p4 = reinterpret_cast<float4*>(g + (size_t)s0 * w);
    for (int i = threadIdx.x; i < n4; i += SH_SITES) {
      float* l = lds + row * (w + 1) + 4 * col;
      if (IN) {
        const float4 v = p4[i];
        l[0] = v.x;
        l[1] = v.y;
        l[2] = v.z;
        l[3] = v.w;
      } else {
        p4[i] = make_float4(l[0], l[1], l[2], l[3]);
      }
      row += dq;
      col += dr;
      if (col >= w4) {
        col -= w4;
        row++;
      }
    }
    return;
  }
  const int n = nb * w;
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

__device__ __forceinline__ void block_in(float* lds, const float* g, int s0, int nb, int w, bool vec) {
  block_copy<true>(lds, const_cast<float*>(g), s0, nb, w, vec);
}

__device__ __forceinline__ void block_out(float* g, float* lds, int s0, int nb, int w, bool vec) {
  block_copy<false>(lds, g, s0, nb, w, vec);
}

// the 48 coefficients of site s (16-byte loads when the array is aligned)
__device__ __forceinline__ void load_sh(const float* sh, int s, float* c, bool vec) {
  if (vec) {
    const float4* src = reinterpret_cast<const float4*>(sh + (size_t)s * 48);
#pragma unroll
    for (int i = 0; i < 12; i++) {
      const float4 v = src[i];
      c[4 * i] = v.x;
      c[4 * i + 1] = v.y;
      c[4 * i + 2] = v.z;
      c[4 * i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 48; i++) c[i] = sh[(size_t)s * 48 + i];
  }
}

__device__ __forceinline__ f3 lds_dir(const float* row, int t, f3 eye) {
  const float ts[3] = {row[3 * t], row[3 * t + 1], row[3 * t + 2]};
  return texel_dir(ts, eye);
}

__global__ void __launch_bounds__(SH_SITES)
    sh_fwd_kernel(const float* __restrict__ sh, const float* __restrict__ texel_sites, int n, int T, f3 eye,
                  float* __restrict__ rgb, bool vec, bool vec_sh) {
  __shared__ float s_io[SH_SITES * SH_ROW];
  const int s0 = blockIdx.x * SH_SITES, nb = min(SH_SITES, n - s0);
  block_in(s_io, texel_sites, s0, nb, 3 * T, vec);
  __syncthreads();
  if ((int)threadIdx.x < nb) {
    float c[48];
    load_sh(sh, s0 + threadIdx.x, c, vec_sh);
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
  block_out(rgb, s_io, s0, nb, 3 * T, vec);
}

__global__ void __launch_bounds__(SH_SITES)
    sh_bwd_kernel(const float* __restrict__ sh, const float* __restrict__ texel_sites,
                  const float* __restrict__ grad_rgb, int n, int T, f3 eye, float* __restrict__ grad_sh, bool vec,
                  bool vec_sh) {
  // texel positions and colour gradients in, then (after a barrier) the
  // sites' 48 coefficient gradients out through the same LDS (rows of 49
  // words), so both directions move whole lines
  static_assert(2 * SH_ROW >= 49, "gradient rows fit the staging space");
  __shared__ float s_mem[2 * SH_SITES * SH_ROW];
  float* s_ts = s_mem;
  float* s_gr = s_mem + SH_SITES * SH_ROW;
  const int s0 = blockIdx.x * SH_SITES, nb = min(SH_SITES, n - s0);
  block_in(s_ts, texel_sites, s0, nb, 3 * T, vec);
  block_in(s_gr, grad_rgb, s0, nb, 3 * T, vec);
  __syncthreads();
  float g[48];
#pragma unroll
  for (int k = 0; k < 48; k++) g[k] = 0.0f;
  if ((int)threadIdx.x < nb) {
    const int s = s0 + threadIdx.x;
    float c[48];
    load_sh(sh, s, c, vec_sh);
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
  }
  __syncthreads();  // the texel rows are read
  if ((int)threadIdx.x < nb) {
#pragma unroll
    for (int k = 0; k < 48; k++) s_mem[threadIdx.x * 49 + k] = g[k];
  }
  __syncthreads();
  block_out(grad_sh, s_mem, s0, nb, 48, vec_sh);
}

}  // namespace pf

using namespace pf;

// 16-byte aligned arrays move through LDS in 16-byte pieces
static bool aligned16(std::initializer_list<const float*> ps) {
  for (const float* q : ps)
    if (reinterpret_cast<uintptr_t>(q) & 15) return false;
  return true;
}

extern "C" {

int pf_sh_fwd(const float* sh, const float* texel_sites, int32_t n_sites, int32_t n_texels, const float* eye,
              float* texel_rgb, void* stream) {
  if (n_sites < 0 || n_texels < 1 || n_texels > PF_MAX_TEXELS || !eye)
    return set_error(PF_ERR_ARG, "pf_sh_fwd: bad arguments");
  if (n_sites == 0) return PF_OK;
  if (!sh || !texel_sites || !texel_rgb) return set_error(PF_ERR_ARG, "pf_sh_fwd: null array");
  sh_fwd_kernel<<<(n_sites + SH_SITES - 1) / SH_SITES, SH_SITES, 0, as_stream(stream)>>>(sh, texel_sites, n_sites, n_texels,
                                                                      f3{eye[0], eye[1], eye[2]}, texel_rgb,
                                                                      aligned16({texel_sites, texel_rgb}), aligned16({sh}));
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
                                                                      n_texels, f3{eye[0], eye[1], eye[2]}, grad_sh,
                                                                      aligned16({texel_sites, grad_texel_rgb}), aligned16({sh, grad_sh}));
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
