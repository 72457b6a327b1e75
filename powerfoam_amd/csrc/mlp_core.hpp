// The learned colour head's per-row arithmetic on BF16 matrix cores, shared by
// the head's own kernels (mlp.hip: texel rows for a camera, training pairs)
// and the compositing forward that evaluates the head in place of a texel
// colour input (raster.hip). Both run the same instruction sequence on the
// same operands, so a texel's colour is bit-identical either way.
//
// For a row of site s seen along unit direction d:
//   x   = [ z_s (16, bf16) ; e(d) (16) ]
//   h1  = relu(W1 x + b1)        W1 [64,32]
//   h2  = relu(W2 h1 + b2)       W2 [64,64]
//   rgb = max(0, W3 h2 + b3)     W3 [3,64]
// with e(d) = (1, Y_1(d) .. Y_15(d)) (the SH head's basis) or the frequency
// encoding (1, d, sin pi d, cos pi d, sin 2 pi d, cos 2 pi d), per component.
// x, h1 and h2 are rounded to bf16 (nearest-even) before each product;
// products accumulate in FP32 (v_mfma_f32_16x16x32_bf16); biases are FP32 and
// enter as the accumulator.
//
// Layers are computed transposed, H^T = W X^T: each weight matrix is the A
// operand (16 output units x 32 inputs), 16 rows of activations the B
// operand (32 inputs x 16 rows). For the 16x16x32 instruction lane
// l = 16g + c16 holds A[c16][8g+j], B[8g+j][c16] (j = 0..7) and D[4g+j][c16]
// (j = 0..3). A D tile holds, per lane, units 16t + 4g + j of row c16; two
// such tiles (t = 2q, 2q + 1) packed to bf16 are exactly a B operand over the
// K-slot order
//   slot 8g + j  <->  unit 32q + 4g + j         (j < 4)
//                     unit 32q + 16 + 4g + j-4  (j >= 4),
// so the next layer's weights are read with their columns in that order and
// no activation leaves the registers.
#pragma once

#include "pf_common.hpp"
#include "sh_basis.hpp"

namespace pf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

constexpr int ML_LAT = 16, ML_IN = 32, ML_HID = 64;

__device__ __forceinline__ uint16_t to_bf(float x) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x)); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 zero4() { return f32x4{0.0f, 0.0f, 0.0f, 0.0f}; }

__device__ __forceinline__ uint32_t pk_bf(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// relu of two packed bf16: a bf16 is negative exactly when its bits are as int16
__device__ __forceinline__ uint32_t relu_pk(uint32_t v) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), (s16x2){0, 0}));
}

// bf16(relu(z)) of two D tiles as one B operand (slot order above)
__device__ __forceinline__ bf16x8 act_op(f32x4 lo, f32x4 hi) {
  const uint4 u = {relu_pk(pk_bf(lo[0], lo[1])), relu_pk(pk_bf(lo[2], lo[3])), relu_pk(pk_bf(hi[0], hi[1])),
                   relu_pk(pk_bf(hi[2], hi[3]))};
  return __builtin_bit_cast(bf16x8, u);
}

// K slot 8g + j of chunk q <-> unit 32q + 16(j / 4) + 4g + j % 4
__device__ __forceinline__ int slot_unit(int q, int g, int j) { return 32 * q + 16 * (j >> 2) + 4 * g + (j & 3); }

// A operand of a 64-column weight row in slot order: columns 32q + 4g .. +3
// and 32q + 16 + 4g .. +3 of row `row`.
__device__ __forceinline__ bf16x8 ld_wrow(const uint16_t* w, int row, int q, int g) {
  const uint2 a = *reinterpret_cast<const uint2*>(w + row * ML_HID + 32 * q + 4 * g);
  const uint2 b = *reinterpret_cast<const uint2*>(w + row * ML_HID + 32 * q + 16 + 4 * g);
  return __builtin_bit_cast(bf16x8, make_uint4(a.x, a.y, b.x, b.y));
}

// The 15 non-constant direction features y[0..14] of e(d) = (1, y).
__device__ __forceinline__ void mlp_encode(f3 d, int encoding, float* y) {
  if (encoding == PF_MLP_ENC_FREQ) {
    // (d, sin pi d, cos pi d, sin 2 pi d, cos 2 pi d) after the constant, from
    // the hardware v_sin_f32 / v_cos_f32 (argument in turns: sin(2 pi x)),
    // whose error (~1e-6 on |x| <= 1) is far below the bf16 rounding that
    // follows; |d| <= 1 keeps the argument in their exact range
    const float dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float half = 0.5f * dd[k];
      y[k] = dd[k];
      y[3 + k] = __builtin_amdgcn_sinf(half);
      y[6 + k] = __builtin_amdgcn_cosf(half);
      y[9 + k] = __builtin_amdgcn_sinf(dd[k]);
      y[12 + k] = __builtin_amdgcn_cosf(dd[k]);
    }
  } else {
    sh_basis(d, y);
  }
}

// Input features 8h .. 8h + 7 of x = [latent ; 1, y] for h = 2, 3 (the
// direction half), packed to bf16.
__device__ __forceinline__ uint4 enc_half(const float* y, int h) {
  return h == 2 ? make_uint4(pk_bf(1.0f, y[0]), pk_bf(y[1], y[2]), pk_bf(y[3], y[4]), pk_bf(y[5], y[6]))
                : make_uint4(pk_bf(y[7], y[8]), pk_bf(y[9], y[10]), pk_bf(y[11], y[12]), pk_bf(y[13], y[14]));
}

}  // namespace pf
