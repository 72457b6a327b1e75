// Scene glue: the rasterizer's per-site inputs from the trainable parameters
// (scene.py:239-337), forward and hand-written backward, one thread per site.
//   radii   = softplus(radii_raw, beta = 100)         scene.py:287-288
//   density = softplus(density_raw, beta = 100)       scene.py:239-240
//   q^ = q / |q|; normal, tangent, bitangent = the columns of R(q^), each
//   normalised                                          scene.py:242-285
//   texel_sites[k] = p + r (s_k0 t + s_k1 b)           scene.py:321-329
//   texel_height[k] = h_k r                             scene.py:337
// softplus follows torch: x when beta x > 20, else log1p(exp(beta x)) / beta.
#include "pf_common.hpp"

namespace pf {

constexpr float SP_BETA = 100.0f;
constexpr float SP_THRESHOLD = 20.0f;

__device__ __forceinline__ float softplus(float x) {
  const float bx = SP_BETA * x;
  return bx > SP_THRESHOLD ? x : log1pf(expf(bx)) / SP_BETA;
}

// d softplus / dx = sigmoid(beta x), 1 above the threshold
__device__ __forceinline__ float softplus_grad(float x) {
  const float bx = SP_BETA * x;
  return bx > SP_THRESHOLD ? 1.0f : 1.0f / (1.0f + expf(-bx));
}

struct Frame {
  float q[4];     // normalised (w, x, y, z)
  float qn;       // |q|
  f3 u[3];        // un-normalised normal, tangent, bitangent
  float len[3];   // their lengths
};

__device__ __forceinline__ Frame make_frame(float4 q4) {
  Frame f;
  f.qn = sqrtf(q4.x * q4.x + q4.y * q4.y + q4.z * q4.z + q4.w * q4.w);
  const float w = q4.x / f.qn, x = q4.y / f.qn, y = q4.z / f.qn, z = q4.w / f.qn;
  f.q[0] = w; f.q[1] = x; f.q[2] = y; f.q[3] = z;
  f.u[0] = f3{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - z * w), 2.0f * (x * z + y * w)};
  f.u[1] = f3{2.0f * (x * y + z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - x * w)};
  f.u[2] = f3{2.0f * (x * z - y * w), 2.0f * (y * z + x * w), 1.0f - 2.0f * (x * x + y * y)};
#pragma unroll
  for (int i = 0; i < 3; i++) f.len[i] = sqrtf(dot(f.u[i], f.u[i]));
  return f;
}

__global__ void __launch_bounds__(256)
    glue_fwd_kernel(int n, int T, const float* __restrict__ points, const float* __restrict__ radii_raw,
                    const float* __restrict__ density_raw, const float4* __restrict__ quats,
                    const float2* __restrict__ offs, const float* __restrict__ height_raw,
                    float* __restrict__ radii, float* __restrict__ density, float* __restrict__ normals,
                    float* __restrict__ texel_sites, float* __restrict__ texel_height) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float r = softplus(radii_raw[s]);
  radii[s] = r;
  density[s] = softplus(density_raw[s]);
  const Frame f = make_frame(quats[s]);
  const f3 nrm = f.u[0] * (1.0f / f.len[0]), tan = f.u[1] * (1.0f / f.len[1]), bit = f.u[2] * (1.0f / f.len[2]);
  normals[3 * (size_t)s + 0] = nrm.x;
  normals[3 * (size_t)s + 1] = nrm.y;
  normals[3 * (size_t)s + 2] = nrm.z;
  const f3 p = ld3(points + 3 * (size_t)s);
  for (int k = 0; k < T; k++) {
    const size_t i = (size_t)s * T + k;
    const float2 o = offs[i];
    const f3 x = p + tan * (o.x * r) + bit * (o.y * r);
    texel_sites[3 * i + 0] = x.x;
    texel_sites[3 * i + 1] = x.y;
    texel_sites[3 * i + 2] = x.z;
    texel_height[i] = height_raw[i] * r;
  }
}

// Adjoint of glue_fwd_kernel. Incoming gradients may be NULL (zero); the
// outputs are written, not accumulated. grad_points gets the texel sites'
// share only (the caller's other uses of the points add theirs).
__global__ void __launch_bounds__(256)
    glue_bwd_kernel(int n, int T, const float* __restrict__ radii_raw, const float* __restrict__ density_raw,
                    const float4* __restrict__ quats, const float2* __restrict__ offs,
                    const float* __restrict__ height_raw, const float* __restrict__ g_radii,
                    const float* __restrict__ g_density, const float* __restrict__ g_normals,
                    const float* __restrict__ g_sites, const float* __restrict__ g_height,
                    float* __restrict__ gp_points, float* __restrict__ gp_radii, float* __restrict__ gp_density,
                    float4* __restrict__ gp_quats, float2* __restrict__ gp_offs, float* __restrict__ gp_height) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float r = softplus(radii_raw[s]);
  const Frame f = make_frame(quats[s]);
  const f3 tan = f.u[1] * (1.0f / f.len[1]), bit = f.u[2] * (1.0f / f.len[2]);
  f3 g_p{0.0f, 0.0f, 0.0f}, g_t{0.0f, 0.0f, 0.0f}, g_b{0.0f, 0.0f, 0.0f};
  float g_r = g_radii ? g_radii[s] : 0.0f;
  for (int k = 0; k < T; k++) {
    const size_t i = (size_t)s * T + k;
    const float2 o = offs[i];
    const f3 gx = g_sites ? ld3(g_sites + 3 * i) : f3{0.0f, 0.0f, 0.0f};
    const float gh = g_height ? g_height[i] : 0.0f;
    g_p = g_p + gx;
    const float gxt = dot(gx, tan), gxb = dot(gx, bit);
    g_r += o.x * gxt + o.y * gxb + gh * height_raw[i];
    g_t = g_t + gx * (o.x * r);
    g_b = g_b + gx * (o.y * r);
    gp_offs[i] = make_float2(r * gxt, r * gxb);
    gp_height[i] = gh * r;
  }
  gp_points[3 * (size_t)s + 0] = g_p.x;
  gp_points[3 * (size_t)s + 1] = g_p.y;
  gp_points[3 * (size_t)s + 2] = g_p.z;
  gp_radii[s] = g_r * softplus_grad(radii_raw[s]);
  gp_density[s] = (g_density ? g_density[s] : 0.0f) * softplus_grad(density_raw[s]);
  // through each column's normalisation: g_u = (g_v - v (v . g_v)) / |u|
  const f3 gn = g_normals ? ld3(g_normals + 3 * (size_t)s) : f3{0.0f, 0.0f, 0.0f};
  const f3 gv[3] = {gn, g_t, g_b};
  f3 gu[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const f3 v = f.u[c] * (1.0f / f.len[c]);
    gu[c] = (gv[c] - v * dot(v, gv[c])) * (1.0f / f.len[c]);
  }
  // columns of R(q^) differentiated by (w, x, y, z)
  const float w = f.q[0], x = f.q[1], y = f.q[2], z = f.q[3];
  const f3 dn[4] = {f3{0.0f, -2.0f * z, 2.0f * y}, f3{0.0f, 2.0f * y, 2.0f * z}, f3{-4.0f * y, 2.0f * x, 2.0f * w},
                    f3{-4.0f * z, -2.0f * w, 2.0f * x}};
  const f3 dt[4] = {f3{2.0f * z, 0.0f, -2.0f * x}, f3{2.0f * y, -4.0f * x, -2.0f * w}, f3{2.0f * x, 0.0f, 2.0f * z},
                    f3{2.0f * w, -4.0f * z, 2.0f * y}};
  const f3 db[4] = {f3{-2.0f * y, 2.0f * x, 0.0f}, f3{2.0f * z, 2.0f * w, -4.0f * x},
                    f3{-2.0f * w, 2.0f * z, -4.0f * y}, f3{2.0f * x, 2.0f * y, 0.0f}};
  float gq[4];
#pragma unroll
  for (int j = 0; j < 4; j++) gq[j] = dot(gu[0], dn[j]) + dot(gu[1], dt[j]) + dot(gu[2], db[j]);
  // through q^ = q / |q|
  const float qg = w * gq[0] + x * gq[1] + y * gq[2] + z * gq[3];
  const float inv = 1.0f / f.qn;
  gp_quats[s] = make_float4((gq[0] - w * qg) * inv, (gq[1] - x * qg) * inv, (gq[2] - y * qg) * inv,
                            (gq[3] - z * qg) * inv);
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_glue_fwd(const pf_glue_params* p, int32_t n_sites, int32_t n_texels, const pf_glue_outputs* o, void* stream) {
  if (!p || !o || n_sites < 0 || n_texels < 1 || n_texels > PF_MAX_TEXELS)
    return set_error(PF_ERR_ARG, "pf_glue_fwd: bad arguments");
  if (n_sites == 0) return PF_OK;
  if (!p->points || !p->radii || !p->density || !p->quaternions || !p->texel_offsets || !p->texel_height ||
      !o->radii || !o->density || !o->normals || !o->texel_sites || !o->texel_height)
    return set_error(PF_ERR_ARG, "pf_glue_fwd: a NULL array");
  if (reinterpret_cast<uintptr_t>(p->quaternions) % 16 || reinterpret_cast<uintptr_t>(p->texel_offsets) % 8)
    return set_error(PF_ERR_ARG, "pf_glue_fwd: quaternions must be 16-byte and texel offsets 8-byte aligned");
  glue_fwd_kernel<<<(n_sites + 255) / 256, 256, 0, as_stream(stream)>>>(
      n_sites, n_texels, p->points, p->radii, p->density, reinterpret_cast<const float4*>(p->quaternions),
      reinterpret_cast<const float2*>(p->texel_offsets), p->texel_height, o->radii, o->density, o->normals,
      o->texel_sites, o->texel_height);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_glue_bwd(const pf_glue_params* p, int32_t n_sites, int32_t n_texels, const pf_glue_outputs* g,
                const pf_glue_grads* gp, void* stream) {
  if (!p || !g || !gp || n_sites < 0 || n_texels < 1 || n_texels > PF_MAX_TEXELS)
    return set_error(PF_ERR_ARG, "pf_glue_bwd: bad arguments");
  if (n_sites == 0) return PF_OK;
  if (!p->radii || !p->density || !p->quaternions || !p->texel_offsets || !p->texel_height || !gp->points ||
      !gp->radii || !gp->density || !gp->quaternions || !gp->texel_offsets || !gp->texel_height)
    return set_error(PF_ERR_ARG, "pf_glue_bwd: a NULL array");
  if (reinterpret_cast<uintptr_t>(p->quaternions) % 16 || reinterpret_cast<uintptr_t>(p->texel_offsets) % 8 ||
      reinterpret_cast<uintptr_t>(gp->quaternions) % 16 || reinterpret_cast<uintptr_t>(gp->texel_offsets) % 8)
    return set_error(PF_ERR_ARG, "pf_glue_bwd: quaternion arrays must be 16-byte, offset arrays 8-byte aligned");
  glue_bwd_kernel<<<(n_sites + 255) / 256, 256, 0, as_stream(stream)>>>(
      n_sites, n_texels, p->radii, p->density, reinterpret_cast<const float4*>(p->quaternions),
      reinterpret_cast<const float2*>(p->texel_offsets), p->texel_height, g->radii, g->density, g->normals,
      g->texel_sites, g->texel_height, gp->points, gp->radii, gp->density, reinterpret_cast<float4*>(gp->quaternions),
      reinterpret_cast<float2*>(gp->texel_offsets), gp->texel_height);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
