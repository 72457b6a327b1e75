// Learned colour head: a shared MLP on BF16 matrix cores.
//
// Fills the same slot as the SH head (the reference's per-texel colour
// pre-pass, color_fn.py:54-184); there is no reference counterpart. For a row
// (a texel, or a training pair) of site s seen along unit direction d:
//   x   = [ z_s (16, bf16) ; e(d) (16) ],  e(d) = (1, Y_1(d) .. Y_15(d)),
//   h1  = relu(W1 x + b1)        W1 [64,32]
//   h2  = relu(W2 h1 + b2)       W2 [64,64]
//   rgb = max(0, W3 h2 + b3)     W3 [3,64]
// with Y_k the real SH basis of the SH head, so e(d) is exactly the vector the
// SH head contracts with its 16 coefficients. Weights and latents are bf16; x,
// h1 and h2 are rounded to bf16 (nearest-even) before each product; products
// accumulate in FP32 (v_mfma_f32_16x16x32_bf16); biases are FP32.
//
// MFMA use. The layers are computed transposed, H^T = W X^T, so each weight
// matrix is the A operand (16 output units x 32 inputs, a 16-byte row slice
// per lane from LDS) and the activations are the B operand (32 inputs x 16
// rows). For the 16x16x32 instruction lane l holds A[l%16][8(l/16)+j],
// B[8(l/16)+j][l%16] and D[4(l/16)+j][l%16], j = 0..7 (A, B) or 0..3 (D): a D
// tile gives each lane 4 consecutive units of one row. The forward chains
// layers through a row-major [row][unit] LDS tile (one 8-byte write, one
// 16-byte read per lane and layer). The backward keeps [unit][row] tiles
// instead, because its weight gradients contract over rows (K = 32 rows per
// step) and need 8 consecutive rows per lane.
#include "pf_common.hpp"
#include "sh_basis.hpp"

namespace pf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int ML_LAT = 16, ML_IN = 32, ML_HID = 64;
constexpr int ML_WPB = 4;           // waves per workgroup
constexpr int FW_RS = ML_HID + 8;   // forward [row][unit] stride (bf16): 144 B rows
constexpr int BW_ROWS = 32;         // backward rows per step
constexpr int BW_RS = BW_ROWS + 8;  // backward [unit][row] stride (bf16): 80 B rows

struct MlpParams {
  int n_rows, n_sites, T;
  const float* texel_sites;  // texel mode: row r is texel r, site r / T
  f3 eye;
  const int32_t* pair_sites;  // pair mode (when set): site and direction per row
  const float* pair_dirs;
  const uint16_t* latents;  // [N,16] bf16
  const uint16_t* w1;
  const float* b1;
  const uint16_t* w2;
  const float* b2;
  const uint16_t* w3;
  const float* b3;
  float* rgb;
  const float* grad_rgb;
  float* g_lat;
  float* g_w1;
  float* g_b1;
  float* g_w2;
  float* g_b2;
  float* g_w3;
  float* g_b3;
};

__device__ __forceinline__ uint16_t to_bf(float x) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x)); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 zero4() { return f32x4{0.0f, 0.0f, 0.0f, 0.0f}; }

__device__ __forceinline__ bool row_ok(const MlpParams& p, int row) {
  if (row >= p.n_rows) return false;
  const int site = p.pair_sites ? p.pair_sites[row] : row / p.T;
  return site >= 0 && site < p.n_sites;
}

// Input slice x[row][8g .. 8g+7]: g = 0, 1 latent halves; g = 2, 3 encoding halves.
__device__ __forceinline__ bf16x8 input_slice(const MlpParams& p, int row, int g) {
  bf16x8 v;
  const int site = row >= p.n_rows ? -1 : (p.pair_sites ? p.pair_sites[row] : row / p.T);
  if (site < 0 || site >= p.n_sites) {
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = static_cast<__bf16>(0.0f);
    return v;
  }
  if (g < 2) return ld8(p.latents + (size_t)site * ML_LAT + 8 * g);
  f3 d;
  if (p.pair_sites) {
    d = ld3(p.pair_dirs + 3 * (size_t)row);
    const float l = sqrtf(dot(d, d));
    d = l > 0.0f ? d * (1.0f / l) : f3{0.0f, 0.0f, 0.0f};
  } else {
    d = texel_dir(p.texel_sites + 3 * (size_t)row, p.eye);
  }
  float y[15];
  sh_basis(d, y);
  float e[16];
  e[0] = 1.0f;
#pragma unroll
  for (int k = 0; k < 15; k++) e[k + 1] = y[k];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = static_cast<__bf16>(g == 2 ? e[j] : e[8 + j]);
  return v;
}

// Weights (and, for the backward, their transposes) staged once per workgroup.
struct alignas(16) WeightsLds {
  uint16_t w1[ML_HID * ML_IN];   // [h1][in]
  uint16_t w2[ML_HID * ML_HID];  // [h2][h1]
  uint16_t w3[16 * ML_HID];      // [out][h2], out rows 3..15 zero
  float b1[ML_HID], b2[ML_HID], b3[4];
};

__device__ void stage_weights(WeightsLds& s, const MlpParams& p) {
  for (int i = threadIdx.x; i < ML_HID * ML_IN; i += blockDim.x) s.w1[i] = p.w1[i];
  for (int i = threadIdx.x; i < ML_HID * ML_HID; i += blockDim.x) s.w2[i] = p.w2[i];
  for (int i = threadIdx.x; i < 16 * ML_HID; i += blockDim.x) s.w3[i] = i < 3 * ML_HID ? p.w3[i] : (uint16_t)0;
  for (int i = threadIdx.x; i < ML_HID; i += blockDim.x) {
    s.b1[i] = p.b1[i];
    s.b2[i] = p.b2[i];
  }
  if (threadIdx.x < 4) s.b3[threadIdx.x] = threadIdx.x < 3 ? p.b3[threadIdx.x] : 0.0f;
}

// ---------------------------------------------------------------- forward
__global__ void __launch_bounds__(WAVE* ML_WPB) mlp_fwd_kernel(MlpParams p) {
  __shared__ WeightsLds sw;
  __shared__ alignas(16) uint16_t s_h_all[ML_WPB][2][16 * FW_RS];
  stage_weights(sw, p);
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c16 = lane & 15;
  uint16_t* sh1 = s_h_all[wv][0];
  uint16_t* sh2 = s_h_all[wv][1];
  const int n_tiles = (p.n_rows + 15) / 16;
  for (int tile = blockIdx.x * ML_WPB + wv; tile < n_tiles; tile += gridDim.x * ML_WPB) {
    const int row = tile * 16 + c16;
    const bf16x8 x = input_slice(p, row, g);
    // layer 1: H1^T tile t = W1[16t.., :] x^T
    wave_sync();
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const f32x4 a = mfma(ld8(sw.w1 + (16 * t + c16) * ML_IN + 8 * g), x, zero4());
      uint16_t o[4];
#pragma unroll
      for (int j = 0; j < 4; j++) o[j] = to_bf(fmaxf(a[j] + sw.b1[16 * t + 4 * g + j], 0.0f));
      *reinterpret_cast<uint2*>(sh1 + c16 * FW_RS + 16 * t + 4 * g) =
          make_uint2(o[0] | ((uint32_t)o[1] << 16), o[2] | ((uint32_t)o[3] << 16));
    }
    wave_sync();
    // layer 2
    const bf16x8 h1a = ld8(sh1 + c16 * FW_RS + 8 * g), h1b = ld8(sh1 + c16 * FW_RS + 32 + 8 * g);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      f32x4 a = mfma(ld8(sw.w2 + (16 * u + c16) * ML_HID + 8 * g), h1a, zero4());
      a = mfma(ld8(sw.w2 + (16 * u + c16) * ML_HID + 32 + 8 * g), h1b, a);
      uint16_t o[4];
#pragma unroll
      for (int j = 0; j < 4; j++) o[j] = to_bf(fmaxf(a[j] + sw.b2[16 * u + 4 * g + j], 0.0f));
      *reinterpret_cast<uint2*>(sh2 + c16 * FW_RS + 16 * u + 4 * g) =
          make_uint2(o[0] | ((uint32_t)o[1] << 16), o[2] | ((uint32_t)o[3] << 16));
    }
    wave_sync();
    // layer 3: lanes of group 0 hold the 3 outputs of their row
    f32x4 a = mfma(ld8(sw.w3 + c16 * ML_HID + 8 * g), ld8(sh2 + c16 * FW_RS + 8 * g), zero4());
    a = mfma(ld8(sw.w3 + c16 * ML_HID + 32 + 8 * g), ld8(sh2 + c16 * FW_RS + 32 + 8 * g), a);
    if (g == 0 && row < p.n_rows) {
      // rows naming no valid site get the biases' colour of a zero input
      float* o = p.rgb + 3 * (size_t)row;
      o[0] = fmaxf(a[0] + sw.b3[0], 0.0f);
      o[1] = fmaxf(a[1] + sw.b3[1], 0.0f);
      o[2] = fmaxf(a[2] + sw.b3[2], 0.0f);
    }
  }
}

// --------------------------------------------------------------- backward
struct alignas(16) WeightsTLds {
  uint16_t w3t[ML_HID * 32];     // [h2][out], out 3..31 zero (K padded to 32)
  uint16_t w2t[ML_HID * ML_HID];  // [h1][h2]
  uint16_t w1t[ML_LAT * ML_HID];  // [in < 16][h1]
};

// One wave's [unit][row] tiles for 32 rows.
struct alignas(16) BwdTiles {
  uint16_t x[ML_IN * BW_RS];
  uint16_t h1[ML_HID * BW_RS];
  uint16_t h2[ML_HID * BW_RS];
  uint16_t dout[16 * BW_RS];
  uint16_t dh1[ML_HID * BW_RS];
  uint16_t dh2[ML_HID * BW_RS];
  float dlat[BW_ROWS * ML_LAT];
};

// B operand slice for row `r` (tile-local) from a [unit][row] tile: units k0..k0+7.
__device__ __forceinline__ bf16x8 col8(const uint16_t* t, int k0, int r) {
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = __builtin_bit_cast(__bf16, t[(k0 + j) * BW_RS + r]);
  return v;
}

// D tile (4 consecutive units u0..u0+3 of row r) into a [unit][row] tile.
__device__ __forceinline__ void put4(uint16_t* t, int u0, int r, const float* v) {
#pragma unroll
  for (int j = 0; j < 4; j++) t[(u0 + j) * BW_RS + r] = to_bf(v[j]);
}

__global__ void __launch_bounds__(WAVE* ML_WPB) mlp_bwd_kernel(MlpParams p) {
  __shared__ WeightsLds sw;
  __shared__ WeightsTLds st;
  __shared__ BwdTiles s_all[ML_WPB];
  stage_weights(sw, p);
  for (int i = threadIdx.x; i < ML_HID * 32; i += blockDim.x) {
    const int h2 = i / 32, o = i % 32;
    st.w3t[i] = o < 3 ? p.w3[o * ML_HID + h2] : (uint16_t)0;
  }
  for (int i = threadIdx.x; i < ML_HID * ML_HID; i += blockDim.x) {
    const int h1 = i / ML_HID, h2 = i % ML_HID;
    st.w2t[i] = p.w2[h2 * ML_HID + h1];
  }
  for (int i = threadIdx.x; i < ML_LAT * ML_HID; i += blockDim.x) {
    const int in = i / ML_HID, h1 = i % ML_HID;
    st.w1t[i] = p.w1[h1 * ML_IN + in];
  }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c16 = lane & 15;
  BwdTiles& s = s_all[wv];
  // output rows 3..15 of dout stay zero
  for (int i = lane; i < 16 * BW_RS; i += WAVE) s.dout[i] = 0;
  __syncthreads();

  f32x4 gw1[4][2], gw2[4][4], gw3[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    gw3[a] = zero4();
#pragma unroll
    for (int b = 0; b < 2; b++) gw1[a][b] = zero4();
#pragma unroll
    for (int b = 0; b < 4; b++) gw2[a][b] = zero4();
  }
  float gb1[4][4], gb2[4][4], gb3[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    gb3[a] = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) gb1[a][j] = gb2[a][j] = 0.0f;
  }

  const int n_steps = (p.n_rows + BW_ROWS - 1) / BW_ROWS;
  for (int step = blockIdx.x * ML_WPB + wv; step < n_steps; step += gridDim.x * ML_WPB) {
    const int base = step * BW_ROWS;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
      const int r = 16 * half + c16;  // tile-local row of this lane's column
      const int row = base + r;
      const bf16x8 x = input_slice(p, row, g);
      wave_sync();
      // single-element reads of the bf16 vector returned element 0 for every
      // index with this compiler, so the raw bits are split instead
      const uint4 xw = __builtin_bit_cast(uint4, x);
      const uint32_t xr[4] = {xw.x, xw.y, xw.z, xw.w};
#pragma unroll
      for (int j = 0; j < 8; j++) s.x[(8 * g + j) * BW_RS + r] = (uint16_t)(xr[j >> 1] >> (16 * (j & 1)));
      // recompute layer 1, keeping pre-activations (D layout: units 16t + 4g + j of row r)
      float z1[4][4], z2[4][4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const f32x4 a = mfma(ld8(sw.w1 + (16 * t + c16) * ML_IN + 8 * g), x, zero4());
        float h[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          z1[t][j] = a[j] + sw.b1[16 * t + 4 * g + j];
          h[j] = fmaxf(z1[t][j], 0.0f);
        }
        put4(s.h1, 16 * t + 4 * g, r, h);
      }
      wave_sync();
      const bf16x8 h1a = col8(s.h1, 8 * g, r), h1b = col8(s.h1, 32 + 8 * g, r);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        f32x4 a = mfma(ld8(sw.w2 + (16 * u + c16) * ML_HID + 8 * g), h1a, zero4());
        a = mfma(ld8(sw.w2 + (16 * u + c16) * ML_HID + 32 + 8 * g), h1b, a);
        float h[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          z2[u][j] = a[j] + sw.b2[16 * u + 4 * g + j];
          h[j] = fmaxf(z2[u][j], 0.0f);
        }
        put4(s.h2, 16 * u + 4 * g, r, h);
      }
      wave_sync();
      f32x4 o = mfma(ld8(sw.w3 + c16 * ML_HID + 8 * g), col8(s.h2, 8 * g, r), zero4());
      o = mfma(ld8(sw.w3 + c16 * ML_HID + 32 + 8 * g), col8(s.h2, 32 + 8 * g, r), o);
      // dL/d(pre-clamp output), lanes of group 0 (outputs 0..2 of row r)
      float d_out[3] = {0.0f, 0.0f, 0.0f};
      if (g == 0 && row_ok(p, row)) {
#pragma unroll
        for (int j = 0; j < 3; j++)
          d_out[j] = (o[j] + sw.b3[j] > 0.0f) ? p.grad_rgb[3 * (size_t)row + j] : 0.0f;
      }
      bf16x8 dob;
#pragma unroll
      for (int j = 0; j < 8; j++) dob[j] = static_cast<__bf16>(j < 3 ? d_out[j] : 0.0f);
      if (g == 0) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          s.dout[j * BW_RS + r] = to_bf(d_out[j]);
          gb3[j] += d_out[j];
        }
      }
      // dH2^T = W3^T dOut^T (K = outputs, padded to 32: only group 0 carries data)
      float dz2[4][4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const f32x4 a = mfma(ld8(st.w3t + (16 * u + c16) * 32 + 8 * g), dob, zero4());
#pragma unroll
        for (int j = 0; j < 4; j++) {
          dz2[u][j] = z2[u][j] > 0.0f ? a[j] : 0.0f;
          gb2[u][j] += dz2[u][j];
        }
        put4(s.dh2, 16 * u + 4 * g, r, dz2[u]);
      }
      wave_sync();
      // dH1^T = W2^T dH2^T
      const bf16x8 d2a = col8(s.dh2, 8 * g, r), d2b = col8(s.dh2, 32 + 8 * g, r);
      float dz1[4][4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        f32x4 a = mfma(ld8(st.w2t + (16 * v + c16) * ML_HID + 8 * g), d2a, zero4());
        a = mfma(ld8(st.w2t + (16 * v + c16) * ML_HID + 32 + 8 * g), d2b, a);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          dz1[v][j] = z1[v][j] > 0.0f ? a[j] : 0.0f;
          gb1[v][j] += dz1[v][j];
        }
        put4(s.dh1, 16 * v + 4 * g, r, dz1[v]);
      }
      wave_sync();
      // dX^T (latent inputs only) = W1[:, :16]^T dH1^T
      f32x4 dx = mfma(ld8(st.w1t + c16 * ML_HID + 8 * g), col8(s.dh1, 8 * g, r), zero4());
      dx = mfma(ld8(st.w1t + c16 * ML_HID + 32 + 8 * g), col8(s.dh1, 32 + 8 * g, r), dx);
#pragma unroll
      for (int j = 0; j < 4; j++) s.dlat[r * ML_LAT + 4 * g + j] = dx[j];
    }
    wave_sync();
    // weight gradients over these 32 rows (K = rows): lane l supplies rows 8g..8g+7
    const int k0 = 8 * g;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const bf16x8 ad1 = ld8(s.dh1 + (16 * a + c16) * BW_RS + k0);
      const bf16x8 ad2 = ld8(s.dh2 + (16 * a + c16) * BW_RS + k0);
#pragma unroll
      for (int b = 0; b < 2; b++) gw1[a][b] = mfma(ad1, ld8(s.x + (16 * b + c16) * BW_RS + k0), gw1[a][b]);
#pragma unroll
      for (int b = 0; b < 4; b++) gw2[a][b] = mfma(ad2, ld8(s.h1 + (16 * b + c16) * BW_RS + k0), gw2[a][b]);
      gw3[a] = mfma(ld8(s.dout + c16 * BW_RS + k0), ld8(s.h2 + (16 * a + c16) * BW_RS + k0), gw3[a]);
    }
    // latent gradients: sum the rows of each site, one atomic per (site, unit)
    const int row_end = min(base + BW_ROWS, p.n_rows);
    if (p.pair_sites) {
      for (int i = lane; i < (row_end - base) * ML_LAT; i += WAVE) {
        const int rr = i / ML_LAT, u = i % ML_LAT;
        const int site = p.pair_sites[base + rr];
        if (site >= 0 && site < p.n_sites) atomicAdd(p.g_lat + (size_t)site * ML_LAT + u, s.dlat[rr * ML_LAT + u]);
      }
    } else {
      const int s0 = base / p.T, s1 = (row_end - 1) / p.T;
      for (int i = lane; i < (s1 - s0 + 1) * ML_LAT; i += WAVE) {
        const int site = s0 + i / ML_LAT, u = i % ML_LAT;
        const int ra = max(site * p.T, base), rb = min((site + 1) * p.T, row_end);
        float acc = 0.0f;
        for (int rr = ra; rr < rb; rr++) acc += s.dlat[(rr - base) * ML_LAT + u];
        atomicAdd(p.g_lat + (size_t)site * ML_LAT + u, acc);
      }
    }
  }

  // flush: D layout puts unit (rows of the product) 16a + 4g + j, column 16b + c16
#pragma unroll
  for (int a = 0; a < 4; a++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int m = 16 * a + 4 * g + j;
#pragma unroll
      for (int b = 0; b < 2; b++) atomicAdd(p.g_w1 + m * ML_IN + 16 * b + c16, gw1[a][b][j]);
#pragma unroll
      for (int b = 0; b < 4; b++) atomicAdd(p.g_w2 + m * ML_HID + 16 * b + c16, gw2[a][b][j]);
    }
    // gw3: product rows = outputs 4g + j (only 0..2), columns h2 = 16a + c16
    if (g == 0) {
#pragma unroll
      for (int j = 0; j < 3; j++) atomicAdd(p.g_w3 + j * ML_HID + 16 * a + c16, gw3[a][j]);
    }
  }
  // biases: partial sums per lane over its rows; reduce over the 16 row lanes
#pragma unroll
  for (int a = 0; a < 4; a++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float v1 = gb1[a][j], v2 = gb2[a][j];
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) {
        v1 += __shfl_xor(v1, m, 64);
        v2 += __shfl_xor(v2, m, 64);
      }
      if (c16 == 0) {
        atomicAdd(p.g_b1 + 16 * a + 4 * g + j, v1);
        atomicAdd(p.g_b2 + 16 * a + 4 * g + j, v2);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    float v = gb3[j];
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
    if (lane == 0) atomicAdd(p.g_b3 + j, v);
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
  if ((reinterpret_cast<uintptr_t>(rows->latents) & 15) || (reinterpret_cast<uintptr_t>(w->w1) & 15))
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
  p.latents = rows->latents;
  p.w1 = w->w1;
  p.b1 = w->b1;
  p.w2 = w->w2;
  p.b2 = w->b2;
  p.w3 = w->w3;
  p.b3 = w->b3;
  return PF_OK;
}

static int persistent_grid(int n_units) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
  }
  const int groups = (n_units + ML_WPB - 1) / ML_WPB;
  return max(1, min(groups, 2 * cus));
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
  mlp_fwd_kernel<<<persistent_grid((p.n_rows + 15) / 16), WAVE * ML_WPB, 0, as_stream(stream)>>>(p);
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
  mlp_bwd_kernel<<<persistent_grid((p.n_rows + BW_ROWS - 1) / BW_ROWS), WAVE * ML_WPB, 0, as_stream(stream)>>>(p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
