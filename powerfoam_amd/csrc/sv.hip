// Spherical-Voronoi texel colour (the reference's colour head,
// color_fn.py:54-184): per texel point, with d = normalize(point - eye),
//   w_i = exp(-temp_i |d - axis_i|),  rgb = max(0, sum w_i val_i / sum w_i + 0.5)
// over the sv_dof lobes, or grey 0.5 outside the camera's cone
// (d . forward < fov_cos_cutoff). One thread per texel; the lobe arrays are
// lobe-major ([D, P, c]), so a wave's loads of one lobe are contiguous.
#include "pf_common.hpp"

namespace pf {

struct SvParams {
  int n;
  int dof;
  const float* points;
  f3 eye, fwd;
  float cutoff;
  const float* axis;
  const float* val;
  const float* temp;
};

__device__ __forceinline__ f3 sv_dir(const SvParams& p, int t, float* len) {
  const f3 v = ld3(p.points + 3 * (size_t)t) - p.eye;
  const float l = sqrtf(dot(v, v));
  *len = l;
  return l > 0.0f ? v * (1.0f / l) : f3{0.0f, 0.0f, 0.0f};
}

__global__ void __launch_bounds__(256) sv_fwd_kernel(SvParams p, float* __restrict__ value,
                                                     float* __restrict__ weights_sum) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.n) return;
  float len;
  const f3 d = sv_dir(p, t, &len);
  float* out = value + 3 * (size_t)t;
  if (dot(d, p.fwd) < p.cutoff) {
    out[0] = out[1] = out[2] = 0.5f;
    weights_sum[t] = 0.0f;
    return;
  }
  float ws = 0.0f;
  f3 vs{0.0f, 0.0f, 0.0f};
  for (int i = 0; i < p.dof; i++) {
    const size_t k = (size_t)i * p.n + t;
    const f3 a = ld3(p.axis + 3 * k);
    const f3 e = d - a;
    const float w = expf(-p.temp[k] * sqrtf(dot(e, e)));
    ws += w;
    vs = vs + ld3(p.val + 3 * k) * w;
  }
  weights_sum[t] = ws;
  out[0] = fmaxf(vs.x / ws + 0.5f, 0.0f);
  out[1] = fmaxf(vs.y / ws + 0.5f, 0.0f);
  out[2] = fmaxf(vs.z / ws + 0.5f, 0.0f);
}

// color_fn.py:105-184: gradients of the lobes' axes, values and temperatures
// (none to the points); a lobe whose normalised weight is below 1e-3 gets
// none. Every output entry is written (zero where the reference leaves it).
__global__ void __launch_bounds__(256)
    sv_bwd_kernel(SvParams p, const float* __restrict__ weights_sum, const float* __restrict__ value,
                  const float* __restrict__ grad_value, float* __restrict__ g_axis, float* __restrict__ g_val,
                  float* __restrict__ g_temp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.n) return;
  float len;
  const f3 d = sv_dir(p, t, &len);
  const bool cut = dot(d, p.fwd) < p.cutoff;
  const float ws = weights_sum[t];
  const f3 v = ld3(value + 3 * (size_t)t);
  f3 g = ld3(grad_value + 3 * (size_t)t);
  if (v.x <= 0.0f) g.x = 0.0f;
  if (v.y <= 0.0f) g.y = 0.0f;
  if (v.z <= 0.0f) g.z = 0.0f;
  const float dp_out = dot(g, v - f3{0.5f, 0.5f, 0.5f});
  for (int i = 0; i < p.dof; i++) {
    const size_t k = (size_t)i * p.n + t;
    f3 ga{0.0f, 0.0f, 0.0f}, gv{0.0f, 0.0f, 0.0f};
    float gt = 0.0f;
    if (!cut) {
      const f3 a = ld3(p.axis + 3 * k);
      const float tmp = p.temp[k];
      const f3 e = d - a;
      const float dist = sqrtf(dot(e, e));
      const float w = expf(-tmp * dist);
      if (!(w / ws < 1e-3f)) {
        gv = g * (w / ws);
        const float c = (dot(g, ld3(p.val + 3 * k)) - dp_out) / ws;
        ga = e * (tmp * w / dist * c);
        gt = -dist * w * c;
      }
    }
    g_axis[3 * k + 0] = ga.x;
    g_axis[3 * k + 1] = ga.y;
    g_axis[3 * k + 2] = ga.z;
    g_val[3 * k + 0] = gv.x;
    g_val[3 * k + 1] = gv.y;
    g_val[3 * k + 2] = gv.z;
    g_temp[k] = gt;
  }
}

static int sv_params(const pf_sv_args* a, SvParams& p, const char* who) {
  if (!a || a->n_points < 0 || a->sv_dof < 1) return set_error(PF_ERR_ARG, "%s: bad arguments", who);
  if (a->n_points > 0 && (!a->points || !a->att_sites || !a->att_values || !a->att_temps))
    return set_error(PF_ERR_ARG, "%s: a NULL array", who);
  p.n = a->n_points;
  p.dof = a->sv_dof;
  p.points = a->points;
  p.eye = f3{a->eye[0], a->eye[1], a->eye[2]};
  p.fwd = f3{a->forward[0], a->forward[1], a->forward[2]};
  p.cutoff = a->fov_cos_cutoff;
  p.axis = a->att_sites;
  p.val = a->att_values;
  p.temp = a->att_temps;
  return PF_OK;
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_sv_fwd(const pf_sv_args* a, float* values, float* weights_sum, void* stream) {
  SvParams p;
  const int st = sv_params(a, p, "pf_sv_fwd");
  if (st != PF_OK) return st;
  if (p.n == 0) return PF_OK;
  if (!values || !weights_sum) return set_error(PF_ERR_ARG, "pf_sv_fwd: NULL output");
  sv_fwd_kernel<<<(p.n + 255) / 256, 256, 0, as_stream(stream)>>>(p, values, weights_sum);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_sv_bwd(const pf_sv_args* a, const float* weights_sum, const float* values, const float* grad_values,
              float* grad_att_sites, float* grad_att_values, float* grad_att_temps, void* stream) {
  SvParams p;
  const int st = sv_params(a, p, "pf_sv_bwd");
  if (st != PF_OK) return st;
  if (p.n == 0) return PF_OK;
  if (!weights_sum || !values || !grad_values || !grad_att_sites || !grad_att_values || !grad_att_temps)
    return set_error(PF_ERR_ARG, "pf_sv_bwd: a NULL array");
  sv_bwd_kernel<<<(p.n + 255) / 256, 256, 0, as_stream(stream)>>>(p, weights_sum, values, grad_values, grad_att_sites,
                                                                  grad_att_values, grad_att_temps);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
