// Real spherical-harmonic basis to degree 3 (constants of the reference's
// sh_utils.py:4-49) and the texel view direction of color_fn.py:72, shared by
// the SH and MLP colour heads.
#pragma once

#include "pf_common.hpp"

namespace pf {

constexpr float C1 = 0.4886025119029199f;
constexpr float C2_0 = 1.0925484305920792f, C2_1 = -1.0925484305920792f, C2_2 = 0.31539156525252005f,
                C2_3 = -1.0925484305920792f, C2_4 = 0.5462742152960396f;
constexpr float C3_0 = -0.5900435899266435f, C3_1 = 2.890611442640554f, C3_2 = -0.4570457994644658f,
                C3_3 = 0.3731763325901154f, C3_4 = -0.4570457994644658f, C3_5 = 1.445305721320277f,
                C3_6 = -0.5900435899266435f;

__device__ __forceinline__ void sh_basis(f3 d, float* y) {
  const float x = d.x, yy_ = d.y, z = d.z;
  const float xx = x * x, yy = yy_ * yy_, zz = z * z;
  y[0] = -C1 * yy_;
  y[1] = C1 * z;
  y[2] = -C1 * x;
  y[3] = C2_0 * x * yy_;
  y[4] = C2_1 * yy_ * z;
  y[5] = C2_2 * (2.0f * zz - xx - yy);
  y[6] = C2_3 * x * z;
  y[7] = C2_4 * (xx - yy);
  y[8] = C3_0 * yy_ * (3.0f * xx - yy);
  y[9] = C3_1 * x * yy_ * z;
  y[10] = C3_2 * yy_ * (4.0f * zz - xx - yy);
  y[11] = C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  y[12] = C3_4 * x * (4.0f * zz - xx - yy);
  y[13] = C3_5 * z * (xx - yy);
  y[14] = C3_6 * x * (xx - 3.0f * yy);
}

__device__ __forceinline__ f3 texel_dir(const float* ts, f3 eye) {
  f3 d = f3{ts[0] - eye.x, ts[1] - eye.y, ts[2] - eye.z};
  const float l = sqrtf(dot(d, d));
  return l > 0.0f ? d * (1.0f / l) : f3{0.0f, 0.0f, 0.0f};
}

}  // namespace pf
