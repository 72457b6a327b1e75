// Learned colour head: a shared MLP on BF16 matrix cores.
//
// Fills the same slot as the SH head (the reference's per-texel colour
// pre-pass, color_fn.py:54-184); there is no reference counterpart. The row
// arithmetic (layers, bf16 rounding points, MFMA operand layout and K-slot
// order) is in mlp_core.hpp, shared with the compositing forward that
// evaluates the head in place of a texel colour input. (The 16x16x16
// instruction, used here for transposes, holds A[c16][4g+j], B[4g+j][c16],
// D[4g+j][c16], j = 0..3.)
#include "mlp_core.hpp"

namespace pf {

constexpr int ML_WPB = 4;  // forward waves per workgroup

struct MlpParams {
  int n_rows, n_sites, T;
  const float* texel_sites;  // texel mode: row r is texel r, site r / T
  f3 eye;
  const int32_t* pair_sites;  // pair mode (when set): site and direction per row
  const float* pair_dirs;
  int encoding;               // PF_MLP_ENC_SH or PF_MLP_ENC_FREQ
  const int32_t* site_list;   // texel mode over listed sites (when set): row r is
  const int32_t* site_count;  // texel r % T of site site_list[r / T], r < count * T
  const uint16_t* latents;  // [N,16] bf16
  const uint16_t* w1;
  const float* b1;
  const uint16_t* w2;
  const float* b2;
  const uint16_t* w3;
  const float* b3;
  float* rgb;
  int out_stride;  // floats between output rows (3, or 4 into (rgb, height) records)
  const float* grad_rgb;
  float* g_lat;
  float* g_w1;
  float* g_b1;
  float* g_w2;
  float* g_b2;
  float* g_w3;
  float* g_b3;
};

// Site of row `row` (< n_rows) and its index into the per-row arrays
// (texel_sites, pair_dirs, grad_rgb): the row itself, or for a site list the
// texel row of the listed site.
__device__ __forceinline__ int row_site(const MlpParams& p, int row, size_t& at) {
  if (p.pair_sites) {
    at = row;
    return p.pair_sites[row];
  }
  if (p.site_list) {
    const int s = p.site_list[row / p.T];
    at = (size_t)s * p.T + row % p.T;
    return s;
  }
  at = row;
  return row / p.T;
}

__device__ __forceinline__ bool row_ok(const MlpParams& p, int row) {
  if (row >= p.n_rows) return false;
  size_t at;
  const int site = row_site(p, row, at);
  return site >= 0 && site < p.n_sites;
}

// ---------------------------------------------------------------- forward
// The forward chains the three products in registers (operand layout and
// K-slot order in mlp_core.hpp): no activation goes through LDS. Weights stay
// in registers for the whole persistent launch; biases ride in as the MFMA
// accumulator.
// The inputs of 64 rows are built one row per lane (latent copy, SH encoding)
// and staged in LDS, from where each 16-row tile reads its B operand.
constexpr int FW_ROWS = 64;  // rows per wave step
constexpr int XS = 40;       // staged input row stride (bf16): 80 B, conflict-free 16-byte reads

// Input row x = [latent ; e(d)] of `row` as 32 bf16 at dst (16-byte aligned);
// zero for rows past the end or naming no valid site.
__device__ __forceinline__ void stage_row(const MlpParams& p, int row, uint16_t* dst) {
  const bool in = row < p.n_rows;
  size_t rc;
  int site = row_site(p, in ? row : 0, rc);
  const bool ok = in && site >= 0 && site < p.n_sites;
  site = ok ? site : 0;
  const uint4* lp = reinterpret_cast<const uint4*>(p.latents + (size_t)site * ML_LAT);
  f3 d;
  if (p.pair_sites) {
    d = ld3(p.pair_dirs + 3 * rc);
    const float l = sqrtf(dot(d, d));
    d = l > 0.0f ? d * (1.0f / l) : f3{0.0f, 0.0f, 0.0f};
  } else {
    d = texel_dir(p.texel_sites + 3 * rc, p.eye);
  }
  float y[15];
  mlp_encode(d, p.encoding, y);
  const uint32_t m = ok ? 0xFFFFFFFFu : 0u;
  const uint4 q0 = lp[0], q1 = lp[1];
  uint4* o = reinterpret_cast<uint4*>(dst);
  o[0] = make_uint4(q0.x & m, q0.y & m, q0.z & m, q0.w & m);
  o[1] = make_uint4(q1.x & m, q1.y & m, q1.z & m, q1.w & m);
  const uint4 e2 = enc_half(y, 2), e3 = enc_half(y, 3);
  o[2] = make_uint4(e2.x & m, e2.y & m, e2.z & m, e2.w & m);
  o[3] = make_uint4(e3.x & m, e3.y & m, e3.z & m, e3.w & m);
}

__global__ void __launch_bounds__(WAVE* ML_WPB) mlp_fwd_kernel(MlpParams pin) {
  __shared__ alignas(16) uint16_t s_x_all[ML_WPB][FW_ROWS * XS];
  MlpParams p = pin;
  if (p.site_list) p.n_rows = *p.site_count * p.T;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c16 = lane & 15;
  uint16_t* sx = s_x_all[wv];
  bf16x8 w1[4], w2[4][2], w3[2];
  f32x4 b1[4], b2[4], b3;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    w1[t] = ld8(p.w1 + (16 * t + c16) * ML_IN + 8 * g);
#pragma unroll
    for (int q = 0; q < 2; q++) w2[t][q] = ld_wrow(p.w2, 16 * t + c16, q, g);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      b1[t][j] = p.b1[16 * t + 4 * g + j];
      b2[t][j] = p.b2[16 * t + 4 * g + j];
    }
  }
#pragma unroll
  for (int q = 0; q < 2; q++) w3[q] = c16 < 3 ? ld_wrow(p.w3, c16, q, g) : __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
#pragma unroll
  for (int j = 0; j < 4; j++) b3[j] = (g == 0 && j < 3) ? p.b3[j] : 0.0f;

  const int n_steps = (p.n_rows + FW_ROWS - 1) / FW_ROWS;
  for (int step = blockIdx.x * ML_WPB + wv; step < n_steps; step += gridDim.x * ML_WPB) {
    const int base = step * FW_ROWS;
    wave_sync();
    stage_row(p, base + lane, sx + lane * XS);
    wave_sync();
#pragma unroll
    for (int tau = 0; tau < FW_ROWS / 16; tau++) {
      const int r = 16 * tau + c16;
      const bf16x8 x = ld8(sx + r * XS + 8 * g);
      f32x4 z1[4], z2[4];
#pragma unroll
      for (int t = 0; t < 4; t++) z1[t] = mfma(w1[t], x, b1[t]);
      const bf16x8 h1[2] = {act_op(z1[0], z1[1]), act_op(z1[2], z1[3])};
#pragma unroll
      for (int u = 0; u < 4; u++) z2[u] = mfma(w2[u][1], h1[1], mfma(w2[u][0], h1[0], b2[u]));
      const bf16x8 h2[2] = {act_op(z2[0], z2[1]), act_op(z2[2], z2[3])};
      const f32x4 o = mfma(w3[1], h2[1], mfma(w3[0], h2[0], b3));
      const int row = base + r;
      // lanes of group 0 hold outputs 0..2 of their row; rows naming no valid
      // site get the biases' colour of a zero input
      if (g == 0 && row < p.n_rows) {
        size_t at = row;
        if (p.site_list) row_site(p, row, at);  // the listed site's texel row
        float* out = p.rgb + (size_t)p.out_stride * at;
        out[0] = fmaxf(o[0], 0.0f);
        out[1] = fmaxf(o[1], 0.0f);
        out[2] = fmaxf(o[2], 0.0f);
      }
    }
  }
}

// --------------------------------------------------------------- backward
// One wave step is 32 rows, two 16-row halves. Per half, in registers:
//   forward again (as above), keeping z1, z2 for the relu masks;
//   dOut  = [o > 0] grad_rgb                    (lanes of group 0)
//   dH2^T = W3^T dOut^T,  dZ2 = [z2 > 0] dH2
//   dH1^T = W2^T dZ2^T,   dZ1 = [z1 > 0] dH1
//   dX^T  = W1[:, :16]^T dZ1^T                  (latent inputs only)
// all in the transposed form, with the same K-slot order as the forward.
// The weight gradients dW1 = dZ1^T X, dW2 = dZ2^T H1, dW3 = dOut^T H2
// contract over rows, so each operand needs 8 rows of one unit per lane,
// while a D tile holds 4 units of one row. An MFMA against a 0/1 selection
// matrix transposes a tile exactly (bf16 values times one, summed with
// zeros), giving 4 rows of one unit per lane; the two halves' tiles together
// are the 8 rows, K slot 8g + j being row 4g + j of half 0 for j < 4 and
// row 16 + 4g + j - 4 of half 1 for j >= 4 (half 0's transposed tiles wait
// in LDS). The weight gradients accumulate in registers over the whole
// persistent launch; at the end the workgroup's waves sum them in LDS and
// issue one atomic per weight. Bias gradients sum the same bf16 dZ the
// weight gradients use: layer 1's is dW1's column of the constant input
// e_0 = 1, layers 2 and 3 sum the transposed dZ2 and dOut (v_dot2 with ones).
// The relu masks of the backward are read from the bf16 activations.
// All A operands (weights, their transposes) sit in LDS as per-lane images:
// one conflict-free 16-byte read each.
constexpr int BW_WPB = 8;  // waves per workgroup
constexpr int BW_ROWS = 32;
enum { IM_W1 = 0, IM_W2 = 4, IM_W3 = 12, IM_W3T = 14, IM_W2T = 18, IM_W1T = 26, IM_N = 28 };


// entry i of the A-operand images: tile i / 512, lane (i / 8) % 64, element i % 8
__device__ __forceinline__ uint16_t image_entry(const MlpParams& p, int i) {
  const int tile = i >> 9, lane = (i >> 3) & 63, j = i & 7;
  const int g = lane >> 4, c = lane & 15;
  if (tile < IM_W2) return p.w1[(16 * tile + c) * ML_IN + 8 * g + j];
  if (tile < IM_W3) {
    const int u = (tile - IM_W2) >> 1, q = (tile - IM_W2) & 1;
    return p.w2[(16 * u + c) * ML_HID + slot_unit(q, g, j)];
  }
  if (tile < IM_W3T) return c < 3 ? p.w3[c * ML_HID + slot_unit(tile - IM_W3, g, j)] : (uint16_t)0;
  if (tile < IM_W2T) return (g == 0 && j < 3) ? p.w3[j * ML_HID + 16 * (tile - IM_W3T) + c] : (uint16_t)0;
  if (tile < IM_W1T) {
    const int v = (tile - IM_W2T) >> 1, q = (tile - IM_W2T) & 1;
    return p.w2[slot_unit(q, g, j) * ML_HID + 16 * v + c];
  }
  return p.w1[slot_unit(tile - IM_W1T, g, j) * ML_IN + c];
}

// Half 0's transposed operands wait in LDS for half 1's: [item][lane] pairs.
enum { TI_H1 = 0, TI_H2 = 4, TI_Z2 = 8, TI_Z1 = 12, TI_X = 16, TI_O = 18, TI_N = 19 };

struct alignas(16) BwdShared {
  union {
    uint16_t im[IM_N * 512];             // A-operand images
    float red[ML_HID * ML_IN + ML_HID * ML_HID + 3 * ML_HID + 2 * ML_HID + 4];  // end-of-launch sums
  };
  float b1[ML_HID], b2[ML_HID], b3[4];
  uint16_t x[BW_WPB][BW_ROWS * XS];     // staged input rows
  float dlat[BW_WPB][BW_ROWS * ML_LAT];  // latent gradient rows
  uint2 t0[BW_WPB][TI_N][WAVE];          // half 0's transposed operands
};

__device__ __forceinline__ bf16x8 im_ld(const BwdShared& sm, int tile, int lane) { return ld8(sm.im + tile * 512 + lane * 8); }

__device__ __forceinline__ f32x4 ld_f4(const float* p) { return *reinterpret_cast<const f32x4*>(p); }

// bf16 [h > 0] d of two D tiles as one B operand: the relu masks are read
// from the bf16 activation operand h of the same slots, which is zero exactly
// where z <= 0 (and for positive z below the smallest bf16 normal, where the
// activation itself is zero). Relu'd bf16 halves lie in [0, 0x7FFF], so
// adding 0x7FFF to each (no carry between halves) sets a half's bit 15 exactly
// when it is nonzero; an arithmetic shift spreads that bit over the half.
__device__ __forceinline__ uint32_t nz_mask(uint32_t h) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(s16x2, h + 0x7FFF7FFFu) >> (s16x2){15, 15});
}

__device__ __forceinline__ bf16x8 pack_masked(f32x4 lo, f32x4 hi, bf16x8 hop) {
  const uint4 m = __builtin_bit_cast(uint4, hop);
  return __builtin_bit_cast(bf16x8, make_uint4(pk_bf(lo[0], lo[1]) & nz_mask(m.x), pk_bf(lo[2], lo[3]) & nz_mask(m.y),
                                               pk_bf(hi[0], hi[1]) & nz_mask(m.z), pk_bf(hi[2], hi[3]) & nz_mask(m.w)));
}

// Transpose of a 16x16 bf16 tile given as the D-layout quadruple of this lane
// (2 packed dwords): returns, packed, the 4 rows 4g .. 4g+3 of column c16.
__device__ __forceinline__ uint2 tr16(uint32_t lo, uint32_t hi, s16x4 sel) {
  const s16x4 a = __builtin_bit_cast(s16x4, make_uint2(lo, hi));
  const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, sel, zero4(), 0, 0, 0);
  return make_uint2(pk_bf(d[0], d[1]), pk_bf(d[2], d[3]));
}

// B operand selecting input columns 16b .. 16b + 15: slot 8g + j is one
// where 8g + j = 16b + c16
__device__ __forceinline__ bf16x8 sel_x(int g, int c16, int b) {
  const int j = 16 * b + c16 - 8 * g;  // the slot of this lane that is one, if 0 <= j < 8
  const uint32_t v = (j & 1) ? (0x3F80u << 16) : 0x3F80u;
  return __builtin_bit_cast(bf16x8, make_uint4(j >> 1 == 0 ? v : 0u, j >> 1 == 1 ? v : 0u, j >> 1 == 2 ? v : 0u,
                                               j >> 1 == 3 ? v : 0u));
}

// acc + the sum of 8 bf16 values (exact products with one, FP32 sums)
__device__ __forceinline__ float sum8(bf16x8 v, float acc) {
  const bf16x2 one = {static_cast<__bf16>(1.0f), static_cast<__bf16>(1.0f)};
  const uint4 u = __builtin_bit_cast(uint4, v);
  acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, u.x), one, acc, false);
  acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, u.y), one, acc, false);
  acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, u.z), one, acc, false);
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, u.w), one, acc, false);
}

__device__ __forceinline__ uint4 dw(bf16x8 v) { return __builtin_bit_cast(uint4, v); }

__device__ __forceinline__ bf16x8 unit_op(uint2 h0, uint2 h1) { return __builtin_bit_cast(bf16x8, make_uint4(h0.x, h0.y, h1.x, h1.y)); }

// One half's transposed operands (4 rows of one unit per lane, packed)
struct HalfT {
  uint2 h1[4], h2[4], z2[4], z1[4], x[2], o;
};

__global__ void __launch_bounds__(WAVE* BW_WPB) mlp_bwd_kernel(MlpParams pin) {
  __shared__ BwdShared sm;
  MlpParams p = pin;
  if (p.site_list) p.n_rows = *p.site_count * p.T;
  for (int i = threadIdx.x; i < IM_N * 512; i += blockDim.x) sm.im[i] = image_entry(p, i);
  for (int i = threadIdx.x; i < ML_HID; i += blockDim.x) {
    sm.b1[i] = p.b1[i];
    sm.b2[i] = p.b2[i];
  }
  if (threadIdx.x < 4) sm.b3[threadIdx.x] = threadIdx.x < 3 ? p.b3[threadIdx.x] : 0.0f;
  __syncthreads();

  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c16 = lane & 15;
  uint16_t* sx = sm.x[wv];
  float* sdl = sm.dlat[wv];
  // selection operands: transpose (16x16x16), and input columns 16b.. (16x16x32)
  s16x4 sel_t;
#pragma unroll
  for (int j = 0; j < 4; j++) sel_t[j] = (c16 - 4 * g == j) ? (short)0x3F80 : (short)0;
  const bf16x8 selx[2] = {sel_x(g, c16, 0), sel_x(g, c16, 1)};

  f32x4 gw1[4][2], gw2[4][4], gw3[4];
  // bias gradients of layers 2 and 3 (layer 1's is dW1's column of the
  // constant input e_0 = 1): sums of the transposed bf16 dZ2 / dOut, per lane
  // for unit 16a + c16 (output c16)
  float gbn2[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gbn3 = 0.0f;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    gw3[a] = zero4();
#pragma unroll
    for (int b = 0; b < 2; b++) gw1[a][b] = zero4();
#pragma unroll
    for (int b = 0; b < 4; b++) gw2[a][b] = zero4();
  }

  const int n_steps = (p.n_rows + BW_ROWS - 1) / BW_ROWS;
  for (int step = blockIdx.x * BW_WPB + wv; step < n_steps; step += gridDim.x * BW_WPB) {
    const int base = step * BW_ROWS;
    wave_sync();
    // incoming colour gradients of both halves (lanes of group 0), loaded
    // before the forward recompute so that it hides their latency
    float gr[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    bool okr[2] = {false, false};
    if (g == 0) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int row = base + 16 * h + c16;
        okr[h] = row_ok(p, row);
        if (okr[h]) {
          size_t at;
          row_site(p, row, at);
#pragma unroll
          for (int j = 0; j < 3; j++) gr[h][j] = p.grad_rgb[3 * at + j];
        }
      }
    }
    // A step whose 32 incoming colour gradients are all zero adds exactly
    // nothing (every product it would accumulate carries dOut) and its
    // sites' latent gradients stay the caller's zeros: skip it. Texel rows
    // of sites no ray shades arrive so from the rasterizer's backward.
    if (ballot(gr[0][0] != 0.0f || gr[0][1] != 0.0f || gr[0][2] != 0.0f || gr[1][0] != 0.0f ||
               gr[1][1] != 0.0f || gr[1][2] != 0.0f) == 0)
      continue;
    if (lane < BW_ROWS) stage_row(p, base + lane, sx + lane * XS);
    wave_sync();
    uint2(*t0)[WAVE] = sm.t0[wv];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int r = 16 * h + c16;
      HalfT T;
      const bf16x8 x = ld8(sx + r * XS + 8 * g);
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const f32x4 d = mfma(x, selx[b], zero4());
        T.x[b] = make_uint2(pk_bf(d[0], d[1]), pk_bf(d[2], d[3]));
        if (h == 0) t0[TI_X + b][lane] = T.x[b];
      }
      bf16x8 h1[2], h2[2];
      {
        f32x4 z1[4];
#pragma unroll
        for (int t = 0; t < 4; t++) z1[t] = mfma(im_ld(sm, IM_W1 + t, lane), x, ld_f4(sm.b1 + 16 * t + 4 * g));
        h1[0] = act_op(z1[0], z1[1]);
        h1[1] = act_op(z1[2], z1[3]);
      }
      {
        f32x4 z2[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
          z2[u] = mfma(im_ld(sm, IM_W2 + 2 * u + 1, lane), h1[1],
                       mfma(im_ld(sm, IM_W2 + 2 * u, lane), h1[0], ld_f4(sm.b2 + 16 * u + 4 * g)));
        h2[0] = act_op(z2[0], z2[1]);
        h2[1] = act_op(z2[2], z2[3]);
      }
      const f32x4 o = mfma(im_ld(sm, IM_W3 + 1, lane), h2[1], mfma(im_ld(sm, IM_W3, lane), h2[0], ld_f4(sm.b3)));
      // dL/d(pre-clamp output): lanes of group 0, outputs 0..2 of row r
      float d_out[3] = {0.0f, 0.0f, 0.0f};
      if (okr[h]) {
#pragma unroll
        for (int j = 0; j < 3; j++) d_out[j] = o[j] > 0.0f ? gr[h][j] : 0.0f;
      }
      const uint32_t do01 = pk_bf(d_out[0], d_out[1]), do23 = pk_bf(d_out[2], 0.0f);
      const bf16x8 dob = __builtin_bit_cast(bf16x8, make_uint4(do01, do23, 0u, 0u));
      T.o = tr16(do01, do23, sel_t);
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const uint4 a1 = dw(h1[t >> 1]), a2 = dw(h2[t >> 1]);
        T.h1[t] = (t & 1) ? tr16(a1.z, a1.w, sel_t) : tr16(a1.x, a1.y, sel_t);
        T.h2[t] = (t & 1) ? tr16(a2.z, a2.w, sel_t) : tr16(a2.x, a2.y, sel_t);
      }
      if (h == 0) {
        t0[TI_O][lane] = T.o;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          t0[TI_H1 + t][lane] = T.h1[t];
          t0[TI_H2 + t][lane] = T.h2[t];
        }
      } else {
        const bf16x8 uo = unit_op(t0[TI_O][lane], T.o);
        gbn3 = sum8(uo, gbn3);
#pragma unroll
        for (int a = 0; a < 4; a++) gw3[a] = mfma(uo, unit_op(t0[TI_H2 + a][lane], T.h2[a]), gw3[a]);
      }
      // dZ2 = [z2 > 0] W3^T dOut
      bf16x8 dz2op[2];
      {
        f32x4 dh2[4];
#pragma unroll
        for (int u = 0; u < 4; u++) dh2[u] = mfma(im_ld(sm, IM_W3T + u, lane), dob, zero4());
        dz2op[0] = pack_masked(dh2[0], dh2[1], h2[0]);
        dz2op[1] = pack_masked(dh2[2], dh2[3], h2[1]);
      }
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const uint4 a3 = dw(dz2op[t >> 1]);
        T.z2[t] = (t & 1) ? tr16(a3.z, a3.w, sel_t) : tr16(a3.x, a3.y, sel_t);
        if (h == 0) t0[TI_Z2 + t][lane] = T.z2[t];
      }
      if (h == 1) {
        bf16x8 ub[4];
#pragma unroll
        for (int b = 0; b < 4; b++) ub[b] = unit_op(t0[TI_H1 + b][lane], T.h1[b]);
#pragma unroll
        for (int a = 0; a < 4; a++) {
          const bf16x8 ua = unit_op(t0[TI_Z2 + a][lane], T.z2[a]);
          gbn2[a] = sum8(ua, gbn2[a]);
#pragma unroll
          for (int b = 0; b < 4; b++) gw2[a][b] = mfma(ua, ub[b], gw2[a][b]);
        }
      }
      // dZ1 = [z1 > 0] W2^T dZ2
      bf16x8 dz1op[2];
      {
        f32x4 dh1[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
          dh1[v] = mfma(im_ld(sm, IM_W2T + 2 * v + 1, lane), dz2op[1],
                        mfma(im_ld(sm, IM_W2T + 2 * v, lane), dz2op[0], zero4()));
        dz1op[0] = pack_masked(dh1[0], dh1[1], h1[0]);
        dz1op[1] = pack_masked(dh1[2], dh1[3], h1[1]);
      }
      // dX^T (latent inputs 4g + j of row r)
      const f32x4 dx = mfma(im_ld(sm, IM_W1T + 1, lane), dz1op[1], mfma(im_ld(sm, IM_W1T, lane), dz1op[0], zero4()));
      *reinterpret_cast<f32x4*>(sdl + r * ML_LAT + 4 * g) = dx;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const uint4 a4 = dw(dz1op[t >> 1]);
        T.z1[t] = (t & 1) ? tr16(a4.z, a4.w, sel_t) : tr16(a4.x, a4.y, sel_t);
        if (h == 0) t0[TI_Z1 + t][lane] = T.z1[t];
      }
      if (h == 1) {
        const bf16x8 ux[2] = {unit_op(t0[TI_X][lane], T.x[0]), unit_op(t0[TI_X + 1][lane], T.x[1])};
#pragma unroll
        for (int a = 0; a < 4; a++) {
          const bf16x8 ua = unit_op(t0[TI_Z1 + a][lane], T.z1[a]);
#pragma unroll
          for (int b = 0; b < 2; b++) gw1[a][b] = mfma(ua, ux[b], gw1[a][b]);
        }
      }
    }
    // latent gradients: sum the rows of each site, one atomic per (site, unit)
    wave_sync();
    const int row_end = min(base + BW_ROWS, p.n_rows);
    if (p.pair_sites) {
      for (int i = lane; i < (row_end - base) * ML_LAT; i += WAVE) {
        const int rr = i / ML_LAT, u = i % ML_LAT;
        const int site = p.pair_sites[base + rr];
        if (site >= 0 && site < p.n_sites) atomicAdd(p.g_lat + (size_t)site * ML_LAT + u, sdl[rr * ML_LAT + u]);
      }
    } else {
      // k: a site's position among the rows (its index, or its list slot)
      const int s0 = base / p.T, s1 = (row_end - 1) / p.T;
      for (int i = lane; i < (s1 - s0 + 1) * ML_LAT; i += WAVE) {
        const int k = s0 + i / ML_LAT, u = i % ML_LAT;
        const int ra = max(k * p.T, base), rb = min((k + 1) * p.T, row_end);
        float acc = 0.0f;
        for (int rr = ra; rr < rb; rr++) acc += sdl[(rr - base) * ML_LAT + u];
        // a site whose rows all fall in this step has no other writer
        const bool whole = k * p.T >= base && ((k + 1) * p.T <= row_end || row_end == p.n_rows);
        const int site = p.site_list ? p.site_list[k] : k;
        if (site >= 0 && site < p.n_sites) {
          if (whole)
            p.g_lat[(size_t)site * ML_LAT + u] = acc;
          else
            atomicAdd(p.g_lat + (size_t)site * ML_LAT + u, acc);
        }
      }
    }
  }

  // end of launch: the workgroup's waves sum into LDS, then one atomic per value
  float* red_w1 = sm.red;
  float* red_w2 = red_w1 + ML_HID * ML_IN;
  float* red_w3 = red_w2 + ML_HID * ML_HID;
  float* red_b1 = red_w3 + 3 * ML_HID;
  float* red_b2 = red_b1 + ML_HID;
  float* red_b3 = red_b2 + ML_HID;
  constexpr int N_RED = ML_HID * ML_IN + ML_HID * ML_HID + 3 * ML_HID + 2 * ML_HID + 3;
  __syncthreads();
  for (int i = threadIdx.x; i < N_RED; i += blockDim.x) sm.red[i] = 0.0f;
  __syncthreads();
  // D layout: product row 16a + 4g + j, column 16b + c16
#pragma unroll
  for (int a = 0; a < 4; a++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int m = 16 * a + 4 * g + j;
#pragma unroll
      for (int b = 0; b < 2; b++) atomicAdd(red_w1 + m * ML_IN + 16 * b + c16, gw1[a][b][j]);
#pragma unroll
      for (int b = 0; b < 4; b++) atomicAdd(red_w2 + m * ML_HID + 16 * b + c16, gw2[a][b][j]);
    }
    // gw3: product rows = outputs 4g + j (0..2 used), columns h2 = 16a + c16
    if (g == 0) {
#pragma unroll
      for (int j = 0; j < 3; j++) atomicAdd(red_w3 + j * ML_HID + 16 * a + c16, gw3[a][j]);
    }
  }
  // biases: per-lane sums over its rows; reduce over the 4 lane groups
#pragma unroll
  for (int a = 0; a < 4; a++) {
    float v = gbn2[a];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (g == 0) atomicAdd(red_b2 + 16 * a + c16, v);
  }
  {
    float v = gbn3;
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (g == 0 && c16 < 3) atomicAdd(red_b3 + c16, v);
  }
  __syncthreads();
  if (threadIdx.x < ML_HID) red_b1[threadIdx.x] = red_w1[threadIdx.x * ML_IN + ML_LAT];
  __syncthreads();
  for (int i = threadIdx.x; i < N_RED; i += blockDim.x) {
    const float v = sm.red[i];
    float* dst;
    int k = i;
    if (k < ML_HID * ML_IN) dst = p.g_w1 + k;
    else if ((k -= ML_HID * ML_IN) < ML_HID * ML_HID) dst = p.g_w2 + k;
    else if ((k -= ML_HID * ML_HID) < 3 * ML_HID) dst = p.g_w3 + k;
    else if ((k -= 3 * ML_HID) < ML_HID) dst = p.g_b1 + k;
    else if ((k -= ML_HID) < ML_HID) dst = p.g_b2 + k;
    else dst = p.g_b3 + (k - ML_HID);
    atomicAdd(dst, v);
  }
}

// Slots in a site list for the four waves of a workgroup, `total` each, with
// one atomic on the list's count per workgroup: returns the wave's first
// slot. Every thread of the workgroup calls it. (An atomic per wave, ~5900
// on one address for 3M sites, serialised to ~70 us.)
__device__ __forceinline__ int reserve_slots(int total, int wv, int lane, int32_t* count) {
  __shared__ int s_cnt[4];
  __shared__ int s_base;
  if (lane == 0) s_cnt[wv] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int sum = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    s_base = sum ? atomicAdd(count, sum) : 0;
  }
  __syncthreads();
  int base = s_base;
  for (int w = 0; w < wv; w++) base += s_cnt[w];
  return base;
}

// Sites with a non-zero colour gradient on some texel, appended to `list`
// (the order within the list is arbitrary). A wave takes ACT_CHUNKS chunks
// of 64 consecutive sites: the chunk's 64 * 3T gradient words are read
// coalesced (16-byte pieces, all issued before any is tested, when the
// array is 16-byte aligned), a non-zero word flags its site in LDS, and the
// workgroup reserves the list slots of its 4 * ACT_CHUNKS chunks with one
// atomic (a per-thread scan of a site's 96 bytes had run at ~5 % of HBM
// bandwidth, one word per lane and step at ~16 %).
constexpr int ACT_CHUNKS = 16;
constexpr int ACT_VEC = 3 * PF_MAX_TEXELS / 4;  // 16-byte pieces per lane of a full chunk (3T <= 24 words a site)

__global__ void __launch_bounds__(256)
    active_sites_kernel(const float* __restrict__ grad_rgb, int n_sites, int T, bool vec, int32_t* list,
                        int32_t* count) {
  __shared__ int s_flag[4][WAVE];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s_first = (blockIdx.x * 4 + wv) * WAVE * ACT_CHUNKS;
  int* flag = s_flag[wv];
  const int n3 = 3 * T;
  uint64_t m[ACT_CHUNKS];
  int total = 0;
#pragma unroll
  for (int c = 0; c < ACT_CHUNKS; c++) {
    const int s0 = s_first + c * WAVE;
    const int ns = max(0, min(WAVE, n_sites - s0));
    flag[lane] = 0;
    wave_sync();
    const float* g = grad_rgb + (size_t)s0 * n3;
    const int nd = ns * n3;
    if (vec) {
      // piece q covers words 4q .. 4q + 3 of the chunk, of at most two sites
      // (3T >= 3); a chunk starts 64 * 3T words (a multiple of 4) after the
      // array's 16-byte aligned start
      float4 v[ACT_VEC];
#pragma unroll
      for (int k = 0; k < ACT_VEC; k++) {
        const int q = lane + WAVE * k;
        v[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (4 * q + 4 <= nd) {
          v[k] = reinterpret_cast<const float4*>(g)[q];
        } else if (4 * q < nd) {  // the array's last, partial piece
          v[k].x = g[4 * q];
          if (4 * q + 1 < nd) v[k].y = g[4 * q + 1];
          if (4 * q + 2 < nd) v[k].z = g[4 * q + 2];
        }
      }
#pragma unroll
      for (int k = 0; k < ACT_VEC; k++) {
        const int w0 = 4 * (lane + WAVE * k);
        const int site = w0 / n3, rem = w0 - site * n3;
        const float c4[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (c4[j] != 0.0f) flag[site + (rem + j >= n3 ? 1 : 0)] = 1;
      }
    } else {
      for (int q = lane; q < nd; q += WAVE)
        if (g[q] != 0.0f) flag[q / n3] = 1;
    }
    wave_sync();
    m[c] = ballot(lane < ns && flag[lane] != 0);
    total += __popcll(m[c]);
    wave_sync();  // the flags are read before the next chunk clears them
  }
  int base = reserve_slots(total, wv, lane, count);
#pragma unroll
  for (int c = 0; c < ACT_CHUNKS; c++) {
    if ((m[c] >> lane) & 1) list[base + __popcll(m[c] & ((1ull << lane) - 1))] = s_first + c * WAVE + lane;
    base += __popcll(m[c]);
  }
}

// Sites with a non-zero mask byte, appended to `list` as active_sites_kernel
// does (a wave takes MASK_CHUNKS chunks of 64 sites, one reservation per
// workgroup).
constexpr int MASK_CHUNKS = 32;

__global__ void __launch_bounds__(256)
    mask_sites_kernel(const uint8_t* __restrict__ mask, const uint8_t* __restrict__ mask_and, int n_sites,
                      int32_t* list, int32_t* count) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s_first = (blockIdx.x * 4 + wv) * WAVE * MASK_CHUNKS;
  uint64_t m[MASK_CHUNKS];
  int total = 0;
#pragma unroll
  for (int c = 0; c < MASK_CHUNKS; c++) {
    const int s = s_first + c * WAVE + lane;
    m[c] = ballot(s < n_sites && mask[s] != 0 && (!mask_and || mask_and[s] != 0));
    total += __popcll(m[c]);
  }
  int base = reserve_slots(total, wv, lane, count);
#pragma unroll
  for (int c = 0; c < MASK_CHUNKS; c++) {
    if ((m[c] >> lane) & 1) list[base + __popcll(m[c] & ((1ull << lane) - 1))] = s_first + c * WAVE + lane;
    base += __popcll(m[c]);
  }
}

static int make_params(const pf_mlp_weights* w, const pf_mlp_rows* rows, MlpParams& p, const char* who) {
  if (!w || !rows || !w->w1 || !w->b1 || !w->w2 || !w->b2 || !w->w3 || !w->b3)
    return set_error(PF_ERR_ARG, "%s: missing weights", who);
  if (rows->n_rows < 0 || (rows->n_rows > 0 && !rows->latents)) return set_error(PF_ERR_ARG, "%s: bad rows", who);
  if (rows->pair_sites) {
    if (!rows->pair_dirs) return set_error(PF_ERR_ARG, "%s: pair mode needs pair_dirs", who);
  } else if (rows->n_rows > 0 && (!rows->texel_sites || rows->n_texels < 1)) {
    return set_error(PF_ERR_ARG, "%s: texel mode needs texel_sites and n_texels >= 1", who);
  }
  if ((reinterpret_cast<uintptr_t>(rows->latents) & 15) || (reinterpret_cast<uintptr_t>(w->w1) & 15) ||
      (reinterpret_cast<uintptr_t>(w->w2) & 15) || (reinterpret_cast<uintptr_t>(w->w3) & 15))
    return set_error(PF_ERR_ARG, "%s: latents and weights must be 16-byte aligned", who);
  if (rows->n_sites < 0 || (!rows->pair_sites && (int64_t)rows->n_sites * rows->n_texels < rows->n_rows))
    return set_error(PF_ERR_ARG, "%s: texel rows exceed n_sites * n_texels", who);
  p = MlpParams{};
  p.n_rows = rows->n_rows;
  p.n_sites = rows->n_sites;
  p.T = rows->n_texels;
  p.texel_sites = rows->texel_sites;
  p.eye = f3{rows->eye[0], rows->eye[1], rows->eye[2]};
  p.pair_sites = rows->pair_sites;
  p.pair_dirs = rows->pair_dirs;
  if (rows->site_list && (rows->pair_sites || !rows->site_count || (int64_t)rows->n_sites * rows->n_texels != rows->n_rows))
    return set_error(PF_ERR_ARG, "%s: a site list needs texel mode, its count and n_rows = n_sites * n_texels", who);
  p.site_list = rows->site_list;
  p.site_count = rows->site_count;
  if (rows->out_stride != 0 && rows->out_stride != 3 && rows->out_stride != 4)
    return set_error(PF_ERR_ARG, "%s: out_stride must be 0, 3 or 4", who);
  if (rows->out_stride == 4 && rows->pair_sites)
    return set_error(PF_ERR_ARG, "%s: (rgb, height) output is for texel rows", who);
  p.out_stride = rows->out_stride ? rows->out_stride : 3;
  if (rows->encoding != PF_MLP_ENC_SH && rows->encoding != PF_MLP_ENC_FREQ)
    return set_error(PF_ERR_ARG, "%s: unknown direction encoding %d", who, rows->encoding);
  p.encoding = rows->encoding;
  p.latents = rows->latents;
  p.w1 = w->w1;
  p.b1 = w->b1;
  p.w2 = w->w2;
  p.b2 = w->b2;
  p.w3 = w->w3;
  p.b3 = w->b3;
  return PF_OK;
}

// Workgroups for a persistent launch: as many as are resident at once (from
// the kernel's occupancy), at most one per wave-step of every wave.
template <class K>
static int persistent_grid(K kernel, int threads, int n_steps) {
  int dev = 0, cus = 256, per_cu = 1;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
  }
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) == hipSuccess && occ > 0) per_cu = occ;
  const int waves = threads / WAVE;
  const int groups = (n_steps + waves - 1) / waves;
  return max(1, min(groups, per_cu * cus));
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_mlp_fwd(const pf_mlp_weights* w, const pf_mlp_rows* rows, float* rgb, void* stream) {
  MlpParams p;
  int st = make_params(w, rows, p, "pf_mlp_fwd");
  if (st) return st;
  if (p.n_rows == 0) return PF_OK;
  if (!rgb) return set_error(PF_ERR_ARG, "pf_mlp_fwd: missing output");
  p.rgb = rgb;
  mlp_fwd_kernel<<<persistent_grid(mlp_fwd_kernel, WAVE * ML_WPB, (p.n_rows + FW_ROWS - 1) / FW_ROWS), WAVE * ML_WPB, 0, as_stream(stream)>>>(p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_mlp_bwd(const pf_mlp_weights* w, const pf_mlp_rows* rows, const float* grad_rgb, pf_mlp_grads* g,
               void* stream) {
  MlpParams p;
  int st = make_params(w, rows, p, "pf_mlp_bwd");
  if (st) return st;
  if (p.n_rows == 0) return PF_OK;
  if (!grad_rgb || !g || !g->latents || !g->w1 || !g->b1 || !g->w2 || !g->b2 || !g->w3 || !g->b3)
    return set_error(PF_ERR_ARG, "pf_mlp_bwd: missing gradient buffers");
  p.grad_rgb = grad_rgb;
  p.g_lat = g->latents;
  p.g_w1 = g->w1;
  p.g_b1 = g->b1;
  p.g_w2 = g->w2;
  p.g_b2 = g->b2;
  p.g_w3 = g->w3;
  p.g_b3 = g->b3;
  mlp_bwd_kernel<<<persistent_grid(mlp_bwd_kernel, WAVE * BW_WPB, (p.n_rows + BW_ROWS - 1) / BW_ROWS), WAVE * BW_WPB, 0, as_stream(stream)>>>(p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_mask_sites(const uint8_t* mask, const uint8_t* mask_and, int32_t n_sites, int32_t* site_list,
                  int32_t* site_count, void* stream) {
  if (n_sites < 0 || (n_sites > 0 && (!mask || !site_list || !site_count)))
    return set_error(PF_ERR_ARG, "pf_mask_sites: bad arguments");
  if (n_sites == 0) return PF_OK;
  const int per_block = 4 * WAVE * MASK_CHUNKS;
  const int grid = (int)(((int64_t)n_sites + per_block - 1) / per_block);
  mask_sites_kernel<<<grid, 256, 0, as_stream(stream)>>>(mask, mask_and, n_sites,
                                                                                         site_list, site_count);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_mlp_active_sites(const float* grad_rgb, int32_t n_sites, int32_t n_texels, int32_t* site_list,
                        int32_t* site_count, void* stream) {
  if (n_sites < 0 || n_texels < 1 || n_texels > PF_MAX_TEXELS ||
      (n_sites > 0 && (!grad_rgb || !site_list || !site_count)))
    return set_error(PF_ERR_ARG, "pf_mlp_active_sites: bad arguments");
  if (n_sites == 0) return PF_OK;
  const int per_block = 4 * WAVE * ACT_CHUNKS;
  const int grid = (int)(((int64_t)n_sites + per_block - 1) / per_block);
  active_sites_kernel<<<grid, 256, 0, as_stream(stream)>>>(
      grad_rgb, n_sites, n_texels, (reinterpret_cast<uintptr_t>(grad_rgb) & 15) == 0, site_list, site_count);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
