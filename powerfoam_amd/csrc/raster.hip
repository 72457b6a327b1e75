// Per-tile compositing, forward and backward, for gfx950.
//
//   forward_kernel   rasterize.py:892-1146
//   backward_kernel  rasterize.py:1150-1514
//   texel surface    texture.py:10-186
//   intersections    rendering_math.py:4-147
//
// One wave64 per 8x8 tile (the reference's 64-thread tile block maps 1:1 onto
// a CDNA wavefront); lane l shades pixel (tile_i*8 + l%8, tile_j*8 + l/8).
// Each workgroup is one tile, so its LDS and slot free up as soon as that
// tile is done. List lengths are very uneven (on a 1M-site view the longest
// list is ~12x the mean), so the host passes a longest-first tile order: the
// longest tiles start first and the short ones fill in around them, instead
// of one long tile starting last and setting the kernel's length.
//
// A wave walks its tile's depth-sorted site list in batches of 64: lanes
// fetch 64 site records with one coalesced gather into LDS, and every
// per-site quantity that does not depend on the pixel (face normals and
// offsets, texel offsets) is computed once per site by the wave into LDS
// instead of once per lane. The tile-wide decisions of the reference become
// ballots: stop when no lane has T >= threshold (the tile max of T is below
// it), skip a site no lane hits. Intersections use cancellation-free forms of
// the reference formulas (ray-sphere: h^2 = r^2 - |oc x d|^2; power face:
// t = ((c-o).n + (|n|^2 + r^2 - r_a^2)/2) / (d.n) with n the decoded
// neighbour delta), so FP32 results stay within 1e-4 of the FP64 oracle
// where the reference's FP32 formulas drift by ~1e-3. Reciprocals and
// exponentials use the hardware v_rcp_f32 / v_exp_f32 (1 ulp).
//
// Exact shortcuts: a site whose texel heights are all zero has identical
// height-stage and colour-stage query points, so the height stage is skipped
// (forward) or its weights reused (backward); a lane whose interval is empty
// after the faces cannot contribute, so it does no texel work.
//
// The backward recomputes each segment back-to-front from the saved final
// log-transmittance (rasterize.py:1219, 1372-1375) and reduces every per-site
// gradient across the wave in registers before touching global memory
// (permlane swaps and DPP, wave_sum16): the neighbour-face gradients per
// distinct face, the site and texel gradients in 16-value slices (two texels
// per slice, recomputed in a final pass so no per-texel accumulators stay
// live); one lane per value issues the atomic.
#include <algorithm>
#include <type_traits>

#include "mlp_core.hpp"

namespace pf {

// One tile (wave) per workgroup: a workgroup's LDS and slot are released as
// soon as its tile is done, so long and short tiles do not wait on each other.
#ifndef PF_WPB
#define PF_WPB 1
#endif
constexpr int WPB = PF_WPB;  // tiles (waves) per workgroup

struct RasterCam {
  f3 eye, right, up, fwd;  // fwd = normalize(up x right)
  int W, H, th, tw;
  const float* ray_maps;
};

static RasterCam make_rcam(const pf_camera& c) {
  RasterCam r;
  r.eye = f3{c.eye[0], c.eye[1], c.eye[2]};
  r.right = f3{c.right[0], c.right[1], c.right[2]};
  r.up = f3{c.up[0], c.up[1], c.up[2]};
  // forward computed on the host in double, rounded once
  double ux = c.up[0], uy = c.up[1], uz = c.up[2], rx = c.right[0], ry = c.right[1], rz = c.right[2];
  double fx = uy * rz - uz * ry, fy = uz * rx - ux * rz, fz = ux * ry - uy * rx;
  double fl = sqrt(fx * fx + fy * fy + fz * fz);
  if (fl > 0) {
    fx /= fl; fy /= fl; fz /= fl;
  }
  r.fwd = f3{(float)fx, (float)fy, (float)fz};
  r.W = c.width;
  r.H = c.height;
  r.th = tiles_h(c);
  r.tw = tiles_w(c);
  r.ray_maps = c.ray_maps;
  return r;
}

// Tile of wave `wv` of workgroup `b` (nw waves each): with an order array,
// slot b*nw + wv takes order[slot] (workgroups start in slot order as CUs
// free up); otherwise the XCD-aware natural order (xcd_group). -1: no tile.
__device__ __forceinline__ int wave_tile(const int32_t* order, int n_tiles, int b, int wv, int nw = WPB) {
  if (order) {
    const int slot = b * nw + wv;
    return slot < n_tiles ? order[slot] : -1;
  }
  bool valid;
  const int t = xcd_group(b, (n_tiles + nw - 1) / nw, valid) * nw + wv;
  return valid && t < n_tiles ? t : -1;
}


struct Ray {
  f3 o, d;
  int pi, pj;
  bool oob;
};

template <bool PINHOLE>
__device__ __forceinline__ Ray lane_ray(const RasterCam& cam, int tile, int lane) {
  Ray r;
  const int ti = tile % cam.th, tj = tile / cam.th;
  int i = ti * PF_TILE_WIDTH + (lane & 7);
  int j = tj * PF_TILE_WIDTH + (lane >> 3);
  r.oob = i >= cam.H || j >= cam.W;
  i = min(i, cam.H - 1);
  j = min(j, cam.W - 1);
  r.pi = i;
  r.pj = j;
  f3 d;
  if (PINHOLE) {
    const float x = 2.0f * (float)j / ((float)cam.W - 1.0f) - 1.0f;
    const float y = 1.0f - 2.0f * (float)i / ((float)cam.H - 1.0f);
    d = cam.right * x + cam.up * y + cam.fwd;
    r.o = cam.eye;
  } else {
    const float* rm = cam.ray_maps + ((size_t)i * cam.W + j) * 6;
    r.o = f3{rm[0], rm[1], rm[2]};
    d = f3{rm[3], rm[4], rm[5]};
  }
  r.d = d * __builtin_amdgcn_rsqf(dot(d, d));
  return r;
}

struct SiteRec {
  float4 sph;  // centre, radius
  float4 nsg;  // normal, sigma
  int site, a0, a1, pad;  // pad: the backward's copy of entry_shades
};

__device__ __forceinline__ float uniform_f(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

// A site record read from LDS by every lane is the same in each: saying so
// puts it in scalar registers, so the per-site loops (faces, texels) count
// and branch in scalar code.
__device__ __forceinline__ SiteRec uniform_site(const SiteRec& r) {
  SiteRec u;
  u.sph = make_float4(uniform_f(r.sph.x), uniform_f(r.sph.y), uniform_f(r.sph.z), uniform_f(r.sph.w));
  u.nsg = make_float4(uniform_f(r.nsg.x), uniform_f(r.nsg.y), uniform_f(r.nsg.z), uniform_f(r.nsg.w));
  u.site = __builtin_amdgcn_readfirstlane(r.site);
  u.a0 = __builtin_amdgcn_readfirstlane(r.a0);
  u.a1 = __builtin_amdgcn_readfirstlane(r.a1);
  u.pad = __builtin_amdgcn_readfirstlane(r.pad);
  return u;
}

// Face (n, w) of a site against neighbour e: n = neighbour delta and
// w = (|n|^2 + r^2 - r_a^2)/2, plus (c - eye).n for pinhole cameras; the ray
// crosses it at t = (w + (c - o).n) / (d.n) -- for a pinhole the second term
// is already folded in. With no encoded deltas the neighbour's
// full-precision sphere is used.
template <bool PINHOLE>
__device__ __forceinline__ float4 face_of(const int32_t* adjacency, const uint16_t* diff, const float4* spheres,
                                          int64_t e, float4 sph, f3 ce) {
  f3 n;
  float ra;
  if (diff) {
    const uint2 raw = *reinterpret_cast<const uint2*>(diff + 4 * e);
    n.x = half_bits_to_float((uint16_t)(raw.x & 0xffff));
    n.y = half_bits_to_float((uint16_t)(raw.x >> 16));
    n.z = half_bits_to_float((uint16_t)(raw.y & 0xffff));
    ra = half_bits_to_float((uint16_t)(raw.y >> 16));
  } else {
    const float4 a = spheres[adjacency[e]];
    n = f3{a.x - sph.x, a.y - sph.y, a.z - sph.z};
    ra = a.w;
  }
  float w = 0.5f * (dot(n, n) + sph.w * sph.w - ra * ra);
  if (PINHOLE) w += dot(ce, n);
  return make_float4(n.x, n.y, n.z, w);
}

// Stages faces e0 .. e0 + ec - 1 of a site (lane l < ec takes face e0 + l)
// and pads the chunk to a multiple of four with a face no ray crosses: its
// offset is NaN, so every crossing t computed from it is NaN and no bound
// comparison accepts it. The face loops then read four faces per LDS round
// trip. Returns the padded count.
// s_adj (when set) takes the faces' neighbour ids, loaded beside the faces.
template <bool PINHOLE>
__device__ __forceinline__ int stage_faces(float4* s_face, const int32_t* adjacency, const uint16_t* diff,
                                           const float4* spheres, int e0, int ec, float4 sph, f3 ce, int lane,
                                           int* s_adj = nullptr) {
  const int ec4 = (ec + 3) & ~3;
  if (lane < ec) {
    if (s_adj) s_adj[lane] = adjacency[e0 + lane];
    s_face[lane] = face_of<PINHOLE>(adjacency, diff, spheres, e0 + lane, sph, ce);
  } else if (lane < ec4) {
    s_face[lane] = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));
  }
  return ec4;
}

// Texel records staged per site: (site - centre, height), (rgb, 0). Slots
// T..7 hold a texel so far away that its weight exp(-10 |x - p|^2 / r^2)
// is exactly 0 (the exponent overflows to -inf) while x - p stays finite:
// it adds exact zeros to every sum, so the texel loops run all eight slots
// without a branch per texel and their LDS reads issue together.
constexpr float TEXEL_FAR = 1e18f;

struct TexelLds {
  float4 pos_h;
  float4 rgb;
};

// Loads the site's texels into LDS (lanes < T) and reports whether any
// height is non-zero.
// bfloat16 -> float: the 16 bits are the float's upper half
__device__ __forceinline__ float bf16_to_float(uint32_t h) { return __builtin_bit_cast(float, h << 16); }

// texel_rgbh_bf16 (4 x bfloat16 per texel), when set, replaces texel_rgbh:
// the colour and height features stored in half the bytes. HEAD: the colour
// of texel `lane` is `head_rgb` (the learned head's, evaluated by the wave),
// and (rgb, height) also go to rgbh_out for the backward.
template <bool HEAD = false>
__device__ __forceinline__ bool stage_texels(TexelLds* s_tex, const float* texel_sites, const float4* texel_rgbh,
                                             const uint2* texel_rgbh_bf16, int site, int T, f3 c, int lane,
                                             f3 head_rgb = f3{0.0f, 0.0f, 0.0f}, float4* rgbh_out = nullptr) {
  bool nz = false;
  if (lane < T) {
    const float* ts = texel_sites + ((size_t)site * T + lane) * 3;
    float4 v;
    if (texel_rgbh_bf16) {
      const uint2 h = texel_rgbh_bf16[(size_t)site * T + lane];
      v = make_float4(bf16_to_float(h.x & 0xffffu), bf16_to_float(h.x >> 16), bf16_to_float(h.y & 0xffffu),
                      bf16_to_float(h.y >> 16));
    } else {
      v = texel_rgbh[(size_t)site * T + lane];
    }
    if (HEAD) {
      v = make_float4(head_rgb.x, head_rgb.y, head_rgb.z, v.w);
      rgbh_out[(size_t)site * T + lane] = v;
    }
    s_tex[lane].pos_h = make_float4(ts[0] - c.x, ts[1] - c.y, ts[2] - c.z, v.w);
    s_tex[lane].rgb = make_float4(v.x, v.y, v.z, 0.0f);
    nz = v.w != 0.0f;
  } else if (lane < PF_MAX_TEXELS) {
    s_tex[lane].pos_h = make_float4(TEXEL_FAR, TEXEL_FAR, TEXEL_FAR, 0.0f);
    s_tex[lane].rgb = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  return ballot(nz) != 0;
}

// The backward's texel records for up to 8 list entries at once (the next
// entries of its walk): lane L stages texel L % 8 of the entry at batch
// position `pos` (its lane group's entry; -1: none) into s_win[L], padded as
// stage_texels pads. One gather for 8 entries instead of a round trip per
// entry. Returns the ballot of non-zero heights (bit 8e + i: texel i of
// window entry e).
__device__ __forceinline__ uint64_t stage_texel_window(TexelLds* s_win, const SiteRec* s_site, int pos,
                                                       const float* texel_sites, const float4* texel_rgbh,
                                                       const uint2* texel_rgbh_bf16, int T, int lane) {
  const int i = lane & 7;
  bool nz = false;
  float4 ph = make_float4(TEXEL_FAR, TEXEL_FAR, TEXEL_FAR, 0.0f), cl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (pos >= 0 && i < T) {
    const int site = s_site[pos].site;
    const float4 sph = s_site[pos].sph;
    const float* ts = texel_sites + ((size_t)site * T + i) * 3;
    float4 v;
    if (texel_rgbh_bf16) {
      const uint2 h = texel_rgbh_bf16[(size_t)site * T + i];
      v = make_float4(bf16_to_float(h.x & 0xffffu), bf16_to_float(h.x >> 16), bf16_to_float(h.y & 0xffffu),
                      bf16_to_float(h.y >> 16));
    } else {
      v = texel_rgbh[(size_t)site * T + i];
    }
    ph = make_float4(ts[0] - sph.x, ts[1] - sph.y, ts[2] - sph.z, v.w);
    cl = make_float4(v.x, v.y, v.z, 0.0f);
    nz = v.w != 0.0f;
  }
  s_win[lane].pos_h = ph;
  s_win[lane].rgb = cl;
  return ballot(nz);
}

struct FwdParams {
  RasterCam cam;
  int n_texels, n_tiles, Q;
  const float4* spheres;
  const float4* nsigmas;
  const float* texel_sites;
  const float4* texel_rgbh;
  const uint2* texel_rgbh_bf16;
  const int32_t* adjacency;
  const int32_t* adj_off;
  const int32_t* adj_cnt;  // faces per site the walk reads (null: the whole range)
  const uint16_t* adj_diff;
  const int64_t* offsets;
  const int32_t* sites;
  const int32_t* order;
  const float* depth_q;
  const float* ray_gt;
  float thr, inv_hw;
  int point_err_on, loss_acc;
  int32_t* early;
  float* color;
  float* log_t;
  float* ndist;
  float* normal;
  float* qdepth;
  float* err;
  float* contrib;
  float* point_err;
  // segments (n_segs = 0: one wave per whole list; seg_count: segments in use)
  int n_segs, seg_len;
  const int32_t* seg_count;
  const int32_t* seg_first;
  const int32_t* seg_tile;
  const int32_t* seg_order;
  uint8_t* entry_shades;  // [M] or null: 1 for each walked entry that shades a ray
  // the learned colour head evaluated in place of the texel colours
  const uint16_t *mlp_w1, *mlp_w2, *mlp_w3, *mlp_lat;
  const float *mlp_b1, *mlp_b2, *mlp_b3;
  int mlp_enc;
  float4* rgbh_out;
  float* seg_logt;   // [n_segs, 64] sum of -sigma dt over the segment
  float* seg_final;  // [n_segs, 64] log T where the segment's wave stopped
  uint8_t* seg_active;  // [n_segs, 64] the lane shaded a site of its walk
};

// The work item of a wave: a whole tile list, or segment k of nk of a tile
// (local sites [k*seg_len, min((k+1)*seg_len, total))). Returns false when
// the slot has no work.
struct WorkItem {
  int tile, seg, k, nk, f0;
};

__device__ __forceinline__ bool work_item(int n_segs, const int32_t* seg_count, const int32_t* seg_order,
                                          const int32_t* seg_tile, const int32_t* seg_first, const int32_t* order,
                                          int n_tiles, int b, int wv, WorkItem& w, int nw = WPB) {
  if (n_segs > 0) {
    const int slot = b * nw + wv;
    if (slot >= n_segs || (seg_count && slot >= *seg_count)) return false;
    w.seg = seg_order[slot];
    w.tile = seg_tile[w.seg];
    w.f0 = seg_first[w.tile];
    w.nk = seg_first[w.tile + 1] - w.f0;
    w.k = w.seg - w.f0;
    return true;
  }
  w.tile = wave_tile(order, n_tiles, b, wv, nw);
  w.seg = -1;
  w.k = 0;
  w.nk = 1;
  w.f0 = 0;
  return w.tile >= 0;
}

// ------------------------------------------------ learned head, in place
// HEAD = true (pf_raster_fwd_args.mlp_head): a site that shades its tile gets
// its texel colours from the learned head, evaluated by the wave on the
// matrix cores with the weights in LDS. One 16-row tile holds two sites: the
// T texels of the site that needs its colours (rows 0..T-1) and those of the
// next site of the list (rows 8..8+T-1), whose colours wait in LDS for the
// case that it shades too (most walked sites do), so an evaluation -- 14
// v_mfma_f32_16x16x32_bf16 (4 + 8 + 2) and the encoding of 16 directions --
// usually serves two sites. The operands and instruction sequence are those
// of pf_mlp_fwd (mlp_core.hpp); a row's result depends on its own column
// only, so the colours are bit-identical to its rows. The weights are staged once per workgroup of HEAD_WPB waves:
// W1 rows padded to 80 B and W2/W3 rows in K-slot order padded to 144 B, so
// every A operand is one conflict-free 16-byte read.
constexpr int HEAD_WPB = 4;
constexpr int MW1S = 40, MW2S = 72;  // row strides (bf16)

struct alignas(16) HeadLds {
  uint16_t w1[ML_HID * MW1S];
  uint16_t w2[ML_HID * MW2S];
  uint16_t w3[3 * MW2S];
  float b1[ML_HID], b2[ML_HID], b3[4];
};

template <class P>
__device__ void stage_head(HeadLds& m, const P& p) {
  for (int i = threadIdx.x; i < ML_HID * ML_IN; i += blockDim.x) m.w1[(i / ML_IN) * MW1S + i % ML_IN] = p.mlp_w1[i];
  for (int i = threadIdx.x; i < ML_HID * ML_HID + 3 * ML_HID; i += blockDim.x) {
    const int k = i < ML_HID * ML_HID ? i : i - ML_HID * ML_HID;
    const int r = k / ML_HID, s = k % ML_HID;  // slot s = 32q + 8g + j
    const int unit = slot_unit(s >> 5, (s >> 3) & 3, s & 7);
    if (i < ML_HID * ML_HID)
      m.w2[r * MW2S + s] = p.mlp_w2[r * ML_HID + unit];
    else
      m.w3[r * MW2S + s] = p.mlp_w3[r * ML_HID + unit];
  }
  for (int i = threadIdx.x; i < ML_HID; i += blockDim.x) {
    m.b1[i] = p.mlp_b1[i];
    m.b2[i] = p.mlp_b2[i];
  }
  if (threadIdx.x < 4) m.b3[threadIdx.x] = threadIdx.x < 3 ? p.mlp_b3[threadIdx.x] : 0.0f;
}

// Colours of the T texels of site_a and of site_b (-1: none): lane c16 of
// lane group 0 gets the rgb of texel c16 of site_a (c16 < T) or of texel
// c16 - 8 of site_b (8 <= c16 < 8 + T). Every lane must take part (the matrix
// instructions read all 64).
template <class P>
__device__ __forceinline__ f3 head_texel_rgb(const HeadLds& m, const P& p, f3 eye, int site_a, int site_b, int T,
                                             int lane) {
  const int g = lane >> 4, c16 = lane & 15;
  const int texel = c16 & 7;
  const int site = c16 < 8 ? site_a : site_b;
  const bool row = texel < T && site >= 0;
  const int rs = row ? site : site_a;
  uint4 xv;
  if (g < 2) {
    xv = *reinterpret_cast<const uint4*>(p.mlp_lat + (size_t)rs * ML_LAT + 8 * g);
  } else {
    const f3 d = texel_dir(p.texel_sites + ((size_t)rs * T + (row ? texel : 0)) * 3, eye);
    float y[15];
    mlp_encode(d, p.mlp_enc, y);
    xv = enc_half(y, g);
  }
  const uint32_t mk = row ? 0xFFFFFFFFu : 0u;
  const bf16x8 x = __builtin_bit_cast(bf16x8, make_uint4(xv.x & mk, xv.y & mk, xv.z & mk, xv.w & mk));
  f32x4 z1[4], z2[4];
#pragma unroll
  for (int t = 0; t < 4; t++)
    z1[t] = mfma(ld8(m.w1 + (16 * t + c16) * MW1S + 8 * g), x, *reinterpret_cast<const f32x4*>(m.b1 + 16 * t + 4 * g));
  const bf16x8 h1[2] = {act_op(z1[0], z1[1]), act_op(z1[2], z1[3])};
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const uint16_t* w = m.w2 + (16 * u + c16) * MW2S + 8 * g;
    z2[u] = mfma(ld8(w + 32), h1[1], mfma(ld8(w), h1[0], *reinterpret_cast<const f32x4*>(m.b2 + 16 * u + 4 * g)));
  }
  const bf16x8 h2[2] = {act_op(z2[0], z2[1]), act_op(z2[2], z2[3])};
  const uint16_t* w3 = m.w3 + min(c16, 2) * MW2S + 8 * g;
  const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
  const bf16x8 w30 = c16 < 3 ? ld8(w3) : zero8, w31 = c16 < 3 ? ld8(w3 + 32) : zero8;
  const f32x4 b3 = g == 0 ? *reinterpret_cast<const f32x4*>(m.b3) : zero4();
  const f32x4 o = mfma(w31, h2[1], mfma(w30, h2[0], b3));
  return f3{fmaxf(o[0], 0.0f), fmaxf(o[1], 0.0f), fmaxf(o[2], 0.0f)};
}

// LOGT = true: the segment prepass of a split list. It walks the segment's
// sites without the early stop and without colour, and stores the per-lane
// sum of -sigma dt, from which every later segment of the tile starts. A
// stop inside an earlier segment leaves every lane below the threshold for
// all later sites (T only decreases), so later segments stop at their first
// site and the split reproduces the serial walk. (Ending the prepass once
// every lane's own T is below the threshold was measured slower: long lists
// on these views rarely get that opaque within one segment.)
// Register budget: six waves per SIMD without the head (80 VGPRs; the
// allocator alone lands at 84, five waves, and the kernel runs 3 % slower
// there despite the 4 spilled registers at six), four with it.
template <bool PINHOLE, bool LOGT, bool HEAD = false>
__global__ void __launch_bounds__(WAVE*(HEAD ? HEAD_WPB : WPB)) __attribute__((amdgpu_waves_per_eu(HEAD ? 4 : 6)))
raster_fwd_kernel(FwdParams p) {
  constexpr int NW = HEAD ? HEAD_WPB : WPB;
  __shared__ SiteRec s_site_all[NW][WAVE];
  __shared__ float4 s_face_all[NW][WAVE];
  // texel records: of the site being shaded (HEAD), or of the next 8 list
  // entries, staged in one gather (stage_texel_window)
  __shared__ TexelLds s_tex_all[NW][HEAD ? PF_MAX_TEXELS : WAVE];
  __shared__ typename std::conditional<HEAD, HeadLds, char>::type s_head;

  if constexpr (HEAD) {
    stage_head(s_head, p);
    __syncthreads();
  }
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  // the learned head's colours of the list entry after the last evaluated
  // one (lanes 0..7), and that entry's position in the batch (-1: none)
  __shared__ typename std::conditional<HEAD, float4[NW][PF_MAX_TEXELS], char>::type s_next_all;
  WorkItem wi;
  if (!work_item(p.n_segs, p.seg_count, p.seg_order, p.seg_tile, p.seg_first, p.order, p.n_tiles, blockIdx.x, wv, wi,
                 NW))
    return;
  if (LOGT && wi.k == wi.nk - 1) return;  // no later segment needs it
  const int tile = wi.tile;
  SiteRec* s_site = s_site_all[wv];
  float4* s_face = s_face_all[wv];
  TexelLds* s_tex = s_tex_all[wv];

  const Ray ray = lane_ray<PINHOLE>(p.cam, tile, lane);
  const size_t px = (size_t)ray.pi * p.cam.W + ray.pj;
  const int T = p.n_texels;

  f3 gt{0.0f, 0.0f, 0.0f};
  if (p.ray_gt) gt = ld3(p.ray_gt + 3 * px);

  float log_t = 0.0f, nd = 0.0f, err = 0.0f;
  f3 rgb{0.0f, 0.0f, 0.0f}, nacc{0.0f, 0.0f, 0.0f};
  if (!LOGT)
    for (int j = 0; j < wi.k; j++) log_t += p.seg_logt[(size_t)(wi.f0 + j) * WAVE + lane];
  // quantiles whose depth an earlier segment wrote: T already below them
  int qidx = 0;
  float qcur = 0.0f;
  if (!LOGT && p.Q > 0) {
    qcur = p.depth_q[px * p.Q];
    if (wi.k > 0) {
      const float t0 = fast_exp(log_t);
      while (qidx < p.Q && t0 < qcur) {
        qidx++;
        if (qidx < p.Q) qcur = p.depth_q[px * p.Q + qidx];
      }
    }
  }

  const int64_t start = p.offsets[tile];
  const int total = (int)(p.offsets[tile + 1] - start);
  const int lo = wi.nk > 1 ? wi.k * p.seg_len : 0;
  const int hi = wi.nk > 1 ? min(lo + p.seg_len, total) : total;
  int early = total;
  bool stop = false;
  bool any_active = false;

  for (int base = lo; base < hi && !stop; base += WAVE) {
    const int cnt = min(WAVE, hi - base);
    int next_pos = -1;  // HEAD: entry whose colours wait in s_next
    wave_sync();
    if (lane < cnt) {
      const int s = p.sites[start + base + lane];
      s_site[lane].sph = p.spheres[s];
      s_site[lane].nsg = p.nsigmas[s];
      s_site[lane].site = s;
      s_site[lane].a0 = p.adj_off[s];
      s_site[lane].a1 = p.adj_cnt ? p.adj_off[s] + p.adj_cnt[s] : p.adj_off[s + 1];
    }
    wave_sync();
    uint64_t win_heights = 0;
    for (int kk = 0; kk < cnt; kk++) {
      const float trans = LOGT ? 1.0f : fast_exp(log_t);
      if (!LOGT && ballot(trans >= p.thr) == 0) {
        early = base + kk;
        stop = true;
        break;
      }
      if (!HEAD && (kk & 7) == 0) {
        // texels of entries kk .. kk + 7: lane group e takes entry kk + e
        const int pos = kk + (lane >> 3) < cnt ? kk + (lane >> 3) : -1;
        wave_sync();
        win_heights = stage_texel_window(s_tex, s_site, pos, p.texel_sites, p.texel_rgbh, p.texel_rgbh_bf16, T, lane);
        wave_sync();
      }
      const SiteRec sr = uniform_site(s_site[kk]);
      const f3 c{sr.sph.x, sr.sph.y, sr.sph.z};
      const float r = sr.sph.w;
      const f3 pn{sr.nsg.x, sr.nsg.y, sr.nsg.z};
      const float sigma = sr.nsg.w;
      const f3 oc = ray.o - c;
      const f3 co = -oc;
      float tn, tf;
      const bool hit = sphere_hit(oc, ray.d, r, tn, tf);
      if (ballot(hit) == 0) continue;

      // power faces (rasterize.py:1038-1071), staged 64 at a time
      const f3 ce = c - p.cam.eye;
      for (int e0 = sr.a0; e0 < sr.a1; e0 += WAVE) {
        const int ec = __builtin_amdgcn_readfirstlane(min(WAVE, sr.a1 - e0));
        wave_sync();
        const int ec4 = stage_faces<PINHOLE>(s_face, p.adjacency, p.adj_diff, p.spheres, e0, ec, sr.sph, ce, lane);
        wave_sync();
        // every lane runs the faces (a lane that misses the sphere only
        // updates values nothing reads): the loop stays in scalar control
        for (int e = 0; e < ec4; e += 4) {
          float4 fq[4];
#pragma unroll
          for (int j = 0; j < 4; j++) fq[j] = s_face[e + j];
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const float4 f = fq[j];
            const f3 n{f.x, f.y, f.z};
            const float dp = dot(ray.d, n);
            const float num = PINHOLE ? f.w : f.w + dot(co, n);
            const float t = acc_div(num, dp);
            // dp < 0 as the complement of dp >= 0: one compare (a NaN dp
            // gives a NaN t, which no comparison below accepts)
            const bool pos = dp >= 0.0f;
            tf = (pos && t < tf) ? t : tf;
            tn = (!pos && t > tn) ? t : tn;
          }
        }
      }
      const bool shade = hit && !ray.oob && tn < tf;
      if (ballot(shade) == 0) continue;

      // texel records of this site (texture.py:23-24)
      f3 head_rgb{0.0f, 0.0f, 0.0f};
      if constexpr (HEAD) {
        float4* s_next = s_next_all[wv];
        if (next_pos == kk) {
          if (lane < PF_MAX_TEXELS) {
            const float4 v = s_next[lane];
            head_rgb = f3{v.x, v.y, v.z};
          }
        } else {
          // this entry's texels and the next entry's (in this batch)
          const int site_b = kk + 1 < cnt ? __builtin_amdgcn_readfirstlane(s_site[kk + 1].site) : -1;
          head_rgb = head_texel_rgb(s_head, p, p.cam.eye, sr.site, site_b, T, lane);
          wave_sync();
          if (lane >= 8 && lane < 16) s_next[lane - 8] = make_float4(head_rgb.x, head_rgb.y, head_rgb.z, 0.0f);
          next_pos = site_b >= 0 ? kk + 1 : -1;
        }
      }
      const TexelLds* tex = HEAD ? s_tex : s_tex + 8 * (kk & 7);
      bool heights = ((win_heights >> (8 * (kk & 7))) & 0xffull) != 0;
      if constexpr (HEAD) {
        wave_sync();
        heights = stage_texels<HEAD>(s_tex, p.texel_sites, p.texel_rgbh, p.texel_rgbh_bf16, sr.site, T, c, lane,
                                     head_rgb, p.rgbh_out);
        wave_sync();
      }
      float alpha_t = 0.0f, perr = 0.0f;
      bool shaded = false;  // an interval of positive length (the backward's "active")
      if (shade) {
        // plane_intersection_fwd (texture.py:10-71)
        const float dpn = dot(pn, ray.d);
        const float cn = dot(co, pn);
        const float kexp = -10.0f * fast_rcp(r * r);
        const float ts0 = acc_div(cn, dpn);
        float height = 0.0f;
        if (heights) {
          const float tq0 = dpn >= 0.0f ? tn : sel_max(tn, ts0);
          const f3 x0 = ray.d * tq0 - co;  // query point relative to the centre
          float hsum = 0.0f, ws0 = 0.0f;
#pragma unroll
          for (int i = 0; i < PF_MAX_TEXELS; i++) {
            const float4 ph = tex[i].pos_h;
            const f3 df{x0.x - ph.x, x0.y - ph.y, x0.z - ph.z};
            const float w = fast_exp(dot(df, df) * kexp);
            hsum += w * ph.w;
            ws0 += w;
          }
          height = hsum * fast_rcp(sel_max(ws0, 1e-20f));
        }
        const float ts1 = heights ? acc_div(cn + height, dpn) : ts0;
        if (LOGT) {
          const float tf2 = (dpn >= 0.0f && ts1 < tf) ? ts1 : tf;
          const float tn2 = (dpn < 0.0f && ts1 > tn) ? ts1 : tn;
          if (tf2 - tn2 > 0.0f) log_t += -sigma * (tf2 - tn2);
          continue;
        }
        const float tq = dpn >= 0.0f ? tn : sel_max(tn, ts1);
        const f3 x1 = ray.d * tq - co;
        f3 rs{0.0f, 0.0f, 0.0f};
        float ws = 0.0f;
#pragma unroll
        for (int i = 0; i < PF_MAX_TEXELS; i++) {
          const float4 ph = tex[i].pos_h;
          const float4 cl = tex[i].rgb;
          const f3 df{x1.x - ph.x, x1.y - ph.y, x1.z - ph.z};
          const float w = fast_exp(dot(df, df) * kexp);
          rs = rs + f3{cl.x, cl.y, cl.z} * w;
          ws += w;
        }
        const f3 col = rs * fast_rcp(sel_max(ws, 1e-20f));
        const float tf2 = (dpn >= 0.0f && ts1 < tf) ? ts1 : tf;
        const float tn2 = (dpn < 0.0f && ts1 > tn) ? ts1 : tn;
        const float dt = tf2 - tn2;
        if (dt > 0.0f) {
          any_active = true;
          shaded = true;
          const float dlt = -sigma * dt;
          const float alpha = 1.0f - fast_exp(dlt);
          const float at = alpha * trans;
          if (p.loss_acc) {
            const f3 df = col - gt;
            err += dot(df, df) * at;
          }
          rgb = rgb + col * at;
          log_t += dlt;
          alpha_t = at;
          if (p.point_err_on) {
            const f3 df = col - gt;
            perr = at * (fabsf(df.x) + fabsf(df.y) + fabsf(df.z));
          }
          const float ndv = dpn;
          if (ndv > 0.0f) nd += (ndv * ndv) * at;
          nacc = nacc + pn * at;
          if (p.Q > 0) {
            const float next_t = fast_exp(log_t);
            while (qidx < p.Q && next_t < qcur) {
              p.qdepth[px * p.Q + qidx] = tn2 + __logf(trans * fast_rcp(qcur)) * fast_rcp(sigma);
              qidx++;
              if (qidx < p.Q) qcur = p.depth_q[px * p.Q + qidx];
            }
          }
        }
      }
      if (!LOGT && p.entry_shades) {
        const bool any = ballot(shaded) != 0;  // every lane votes
        if (any && lane == 0) p.entry_shades[start + base + kk] = 1;
      }
      // contrib / point_err: one atomic per site per wave (rasterize.py:1101-1110)
      if (!LOGT && ballot(alpha_t > 0.0f) != 0) {
        const float sa = wave_sum(alpha_t);
        if (lane == 0) atomicAdd(&p.contrib[sr.site], sa * p.inv_hw);
        if (p.point_err_on) {
          const float se = wave_sum(perr);
          if (lane == 0) atomicAdd(&p.point_err[sr.site], se * p.inv_hw);
        }
      }
    }
  }

  if (LOGT) {
    p.seg_logt[(size_t)wi.seg * WAVE + lane] = log_t;
    return;
  }
  if (wi.seg >= 0) p.seg_active[(size_t)wi.seg * WAVE + lane] = any_active ? 1 : 0;
  if (wi.nk == 1) {
    if (lane == 0) p.early[tile] = early;
    if (!ray.oob) {
      p.color[3 * px + 0] = rgb.x;
      p.color[3 * px + 1] = rgb.y;
      p.color[3 * px + 2] = rgb.z;
      p.log_t[px] = log_t;
      p.ndist[px] = nd;
      p.normal[3 * px + 0] = nacc.x;
      p.normal[3 * px + 1] = nacc.y;
      p.normal[3 * px + 2] = nacc.z;
      if (p.loss_acc && p.err) p.err[px] = err + dot(gt, gt) * fast_exp(log_t);
    }
    return;
  }
  // a segment of a split list: partial sums into the zeroed outputs; the
  // final log T and the early-stop index are settled by the finishing kernel
  if (lane == 0 && (stop || wi.k == wi.nk - 1)) atomicMin(&p.early[tile], early);
  p.seg_final[(size_t)wi.seg * WAVE + lane] = log_t;
  if (!ray.oob) {
    atomicAdd(&p.color[3 * px + 0], rgb.x);
    atomicAdd(&p.color[3 * px + 1], rgb.y);
    atomicAdd(&p.color[3 * px + 2], rgb.z);
    atomicAdd(&p.ndist[px], nd);
    atomicAdd(&p.normal[3 * px + 0], nacc.x);
    atomicAdd(&p.normal[3 * px + 1], nacc.y);
    atomicAdd(&p.normal[3 * px + 2], nacc.z);
    if (p.loss_acc && p.err) atomicAdd(&p.err[px], err);
  }
}

// Split lists: log T of a pixel is where its tile's walk stopped -- in the
// segment holding the early-stop index, or the last one.
template <bool PINHOLE>
__global__ void __launch_bounds__(WAVE) raster_fwd_finish_kernel(FwdParams p) {
  const int tile = blockIdx.x, lane = threadIdx.x;
  const int f0 = p.seg_first[tile], nk = p.seg_first[tile + 1] - f0;
  if (nk <= 1) return;
  const int ks = min(p.early[tile] / p.seg_len, nk - 1);
  const float lt = p.seg_final[(size_t)(f0 + ks) * WAVE + lane];
  const Ray ray = lane_ray<PINHOLE>(p.cam, tile, lane);
  if (ray.oob) return;
  const size_t px = (size_t)ray.pi * p.cam.W + ray.pj;
  p.log_t[px] = lt;
  if (p.loss_acc && p.err) {
    const f3 gt = ld3(p.ray_gt + 3 * px);
    atomicAdd(&p.err[px], dot(gt, gt) * fast_exp(lt));
  }
}

__global__ void __launch_bounds__(256) fill_int_kernel(int32_t* x, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// ----------------------------------------------------------------- backward

struct BwdParams {
  RasterCam cam;
  int n_texels, n_tiles, Q, exact;
  const float4* spheres;
  const float4* nsigmas;
  const float* texel_sites;
  const float4* texel_rgbh;
  const uint2* texel_rgbh_bf16;
  const int32_t* adjacency;
  const int32_t* adj_off;
  const int32_t* adj_cnt;  // faces per site the walk reads (null: the whole range)
  const uint16_t* adj_diff;
  const int64_t* offsets;
  const int32_t* sites;
  const int32_t* order;
  const float* depth_q;
  const int32_t* early;
  const float* log_t;
  const float* g_color;
  const float* g_log_t;
  const float* g_ndist;
  const float* g_normal;
  const float* g_qdepth;
  const float* g_contrib;
  float inv_hw;
  float* g_spheres;
  float* g_nsigmas;
  float* g_tsites;
  float* g_trgb;
  float* g_theight;
  // segments (n_segs = 0: one wave per whole list; seg_count: segments in use)
  int n_segs, seg_len;
  const int32_t* seg_count;
  const int32_t* seg_first;
  const int32_t* seg_tile;
  const int32_t* seg_order;
  const float* seg_logt;  // [n_segs, 64] from the forward
  const uint8_t* seg_active;  // [n_segs, 64] from the forward
  float* seg_dsum;        // [n_segs, 64] per-segment sum of dL/dlog T increments
  const uint8_t* entry_shades;  // [M] from the forward, or null: entries marked 0 are skipped
};

constexpr int RED_SLICE = 16;  // values reduced per pass

// Sums 16 values over the wave without LDS: a butterfly reduce-scatter in
// registers. permlane32_swap exchanges the half-waves and permlane16_swap the
// odd and even rows of two values at once, so each exchange halves the values
// a lane carries; DPP row_ror:8 and row_half_mirror pair lanes i, i^8 and
// i, i^7 inside a row; two quad_perm adds finish the sum. Lanes 4u .. 4u+3
// return the sum of v[u]. Every lane must be active.
__device__ __forceinline__ float wave_sum16(const float* v, int lane) {
  float r[8], s[4], t[2];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    // lanes 0-31 keep v[q], lanes 32-63 keep v[q + 8]
    float a = v[q], b = v[q + 8];
    permlane32_swap(a, b);
    r[q] = a + b;
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    // even rows keep r[q], odd rows r[q + 4]
    float a = r[q], b = r[q + 4];
    permlane16_swap(a, b);
    s[q] = a + b;
  }
  const bool b3 = lane & 8, b2 = lane & 4;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    // lanes with bit 3 clear keep s[q], the others s[q + 2]; partner i ^ 8
    const float keep = b3 ? s[q + 2] : s[q], send = b3 ? s[q] : s[q + 2];
    t[q] = keep + dpp<0x128>(send);  // row_ror:8
  }
  // bit 2 clear keeps t[0], set t[1]; partner i ^ 7 (row_half_mirror)
  const float keep = b2 ? t[1] : t[0], send = b2 ? t[0] : t[1];
  float u = keep + dpp<0x141>(send);
  u += dpp<0xB1>(u);  // quad_perm [1,0,3,2]
  u += dpp<0x4E>(u);  // quad_perm [2,3,0,1]
  return u;
}

// Destination of reduced value v of a site (v < 8 + 7T).
__device__ __forceinline__ float* grad_slot(const BwdParams& p, int site, int T, int v) {
  if (v < 4) return p.g_spheres + 4 * (size_t)site + v;
  if (v < 8) return p.g_nsigmas + 4 * (size_t)site + (v - 4);
  const int i = (v - 8) / 7, q = (v - 8) % 7;
  if (q < 3) return p.g_tsites + ((size_t)site * T + i) * 3 + q;
  if (q < 6) return p.g_trgb + ((size_t)site * T + i) * 3 + (q - 3);
  return p.g_theight + (size_t)site * T + i;
}

// The backward's register budget: three waves per SIMD (at most 168
// VGPRs); left to itself the allocator lands just above and the kernel runs
// two waves per SIMD.
#ifndef PF_BWD_WAVES
#define PF_BWD_WAVES 3
#endif
#define PF_BWD_OCC __attribute__((amdgpu_waves_per_eu(PF_BWD_WAVES)))
// DSUM = true: the segment prepass of a split list. dL/dlog T, the adjoint
// that carries information from the back of a list to its front, gains an
// additive increment at every site; this pass sums the increments of one
// segment so that every earlier segment can start from the value the later
// ones leave. The geometry, colour and quantile state are those of the main
// pass; no parameter gradients are formed.
template <bool PINHOLE, bool DSUM>
__global__ void __launch_bounds__(WAVE* WPB) PF_BWD_OCC raster_bwd_kernel(BwdParams p) {
  __shared__ SiteRec s_site_all[WPB][WAVE];
  __shared__ float4 s_face_all[WPB][WAVE];
  __shared__ TexelLds s_win_all[WPB][WAVE];  // texels of 8 walk entries
  __shared__ float s_gc_all[WPB][WAVE];      // dL/dcontrib of the batch's sites
  __shared__ int s_fadj_all[WPB][WAVE];      // neighbour ids of the staged faces

  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  WorkItem wi;
  if (!work_item(p.n_segs, p.seg_count, p.seg_order, p.seg_tile, p.seg_first, p.order, p.n_tiles, blockIdx.x, wv, wi)) return;
  const int tile = wi.tile;
  const int early = p.early[tile];
  // this segment's share of the walked sites [0, early)
  const int lo = wi.nk > 1 ? wi.k * p.seg_len : 0;
  const int hi = wi.nk > 1 ? min(lo + p.seg_len, early) : early;
  if (hi <= lo) return;
  const int k_last = wi.nk > 1 ? (early - 1) / p.seg_len : 0;  // segment of the last walked site
  if (DSUM && wi.k == 0) return;  // no earlier segment needs it
  SiteRec* s_site = s_site_all[wv];
  float4* s_face = s_face_all[wv];
  TexelLds* s_win = s_win_all[wv];
  float* s_gc = s_gc_all[wv];
  int* s_fadj = s_fadj_all[wv];

  const Ray ray = lane_ray<PINHOLE>(p.cam, tile, lane);
  const size_t px = (size_t)ray.pi * p.cam.W + ray.pj;
  const int T = p.n_texels;
  const bool exact = p.exact != 0;
  const float pl_eps = exact ? 0.0f : 1e-6f;
  const float cut = exact ? 0.0f : 1e-3f;
  const float q_eps = exact ? 0.0f : 1e-6f;

  // state after the segment's last site: log T and dL/dlog T (black
  // background: no colour term), and the quantiles later segments handle
  float log_t, dLdlog_t;
  int qidx = p.Q - 1;
  if (wi.k == k_last) {
    log_t = p.log_t[px];
    dLdlog_t = DSUM ? 0.0f : p.g_log_t[px];
  } else {
    log_t = 0.0f;
    for (int j = 0; j <= wi.k; j++) log_t += p.seg_logt[(size_t)(wi.f0 + j) * WAVE + lane];
    dLdlog_t = 0.0f;
    if (!DSUM) {
      dLdlog_t = p.g_log_t[px];
      for (int j = wi.k + 1; j <= k_last; j++) dLdlog_t += p.seg_dsum[(size_t)(wi.f0 + j) * WAVE + lane];
    }
    // later segments handle the quantiles below T at this segment's end --
    // provided this lane shades a site there (a quantile is handled at an
    // active site only)
    bool later = false;
    for (int j = wi.k + 1; j <= k_last; j++) later = later || p.seg_active[(size_t)(wi.f0 + j) * WAVE + lane] != 0;
    if (later) {
      const float t_end = fast_exp(log_t);
      while (qidx >= 0 && t_end > p.depth_q[px * p.Q + qidx]) qidx--;
    }
  }
  const f3 g_rgb = ld3(p.g_color + 3 * px);
  const f3 g_n = ld3(p.g_normal + 3 * px);
  const float g_nd = p.g_ndist[px];
  float qcur = qidx >= 0 ? p.depth_q[px * p.Q + qidx] : 1.0f;

  const int64_t start = p.offsets[tile];

  for (int top = hi; top > lo; top -= WAVE) {
    const int cnt = min(WAVE, top - lo);
    const int base = top - cnt;  // batch covers [base, top)
    wave_sync();
    if (lane < cnt) {
      const int s = p.sites[start + base + lane];
      s_site[lane].sph = p.spheres[s];
      s_site[lane].nsg = p.nsigmas[s];
      s_site[lane].site = s;
      s_site[lane].a0 = p.adj_off[s];
      s_site[lane].a1 = p.adj_cnt ? p.adj_off[s] + p.adj_cnt[s] : p.adj_off[s + 1];
      s_site[lane].pad = p.entry_shades ? p.entry_shades[start + base + lane] : 1;
      s_gc[lane] = p.g_contrib[s];
    }
    // the entries to walk, back to front: an entry that shades no ray of the
    // tile (as the forward found) adds nothing, so its geometry is skipped
    uint64_t rest = ballot(lane < cnt && (!p.entry_shades || p.entry_shades[start + base + lane] != 0));
    wave_sync();
    int walked = 0;
    uint64_t win_heights = 0;
    while (rest) {
      const int kk = 63 - __builtin_clzll(rest);
      const int slot = walked & 7;
      if (slot == 0) {
        // texels of this entry and the next 7 to walk: lane group e takes
        // the e-th remaining entry
        int pos = -1;
        uint64_t m = rest;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int pe = m ? 63 - __builtin_clzll(m) : -1;
          pos = (lane >> 3) == e ? pe : pos;
          m = m ? (m & ~(1ull << pe)) : m;
        }
        wave_sync();
        win_heights = stage_texel_window(s_win, s_site, pos, p.texel_sites, p.texel_rgbh, p.texel_rgbh_bf16, T, lane);
        wave_sync();
      }
      TexelLds* s_tex = s_win + 8 * slot;
      const bool heights = ((win_heights >> (8 * slot)) & 0xffull) != 0;
      walked++;
      rest &= ~(1ull << kk);
      const SiteRec sr = uniform_site(s_site[kk]);
      const f3 c{sr.sph.x, sr.sph.y, sr.sph.z};
      const float r = sr.sph.w;
      const f3 pn{sr.nsg.x, sr.nsg.y, sr.nsg.z};
      const float sigma = sr.nsg.w;
      const f3 oc = ray.o - c;
      const f3 co = -oc;
      float tn, tf;
      const bool hit = sphere_hit(oc, ray.d, r, tn, tf);
      if (ballot(hit) == 0) continue;
      const int n_adj = sr.a1 - sr.a0;
      // id: -2 none, -1 sphere, 0..n-1 face, n plane (rasterize.py:1296)
      int tn_id = tn < 1e-8f ? -2 : -1;
      int tf_id = tf < 1e-8f ? -2 : -1;
      const f3 ce = c - p.cam.eye;
      for (int e0 = sr.a0; e0 < sr.a1; e0 += WAVE) {
        const int ec = __builtin_amdgcn_readfirstlane(min(WAVE, sr.a1 - e0));
        wave_sync();
        const int ec4 = stage_faces<PINHOLE>(s_face, p.adjacency, p.adj_diff, p.spheres, e0, ec, sr.sph, ce, lane,
                                             s_fadj);
        wave_sync();
        // every lane runs the faces (a lane that misses the sphere only
        // updates values nothing reads): the loop stays in scalar control
        for (int e = 0; e < ec4; e += 4) {
          float4 fq[4];
#pragma unroll
          for (int j = 0; j < 4; j++) fq[j] = s_face[e + j];
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const float4 f = fq[j];
            const f3 n{f.x, f.y, f.z};
            const float dp = dot(ray.d, n);
            const float num = PINHOLE ? f.w : f.w + dot(co, n);
            const float t = acc_div(num, dp);
            const int id = e0 - sr.a0 + e + j;
            const bool pos = dp >= 0.0f;  // one compare, as in the forward
            if (pos && t < tf) {
              tf_id = id;
              tf = t;
            }
            if (!pos && t > tn) {
              tn_id = id;
              tn = t;
            }
          }
        }
      }
      const bool shade = hit && !ray.oob && tn < tf;
      if (ballot(shade) == 0) continue;

      // ---- recompute plane_intersection_fwd (texture.py:10-71), keeping the
      // colour-stage weights
      const float tn0 = tn;
      const int tn0_id = tn_id, tf0_id = tf_id;
      const float kexp = -10.0f * fast_rcp(r * r);
      float w1[PF_MAX_TEXELS];
      float dpn = 0.0f, cn = 0.0f, ts0 = 0.0f, ts1 = 0.0f, height = 0.0f, ws0 = 1.0f, ws = 1.0f;
      f3 col{0.0f, 0.0f, 0.0f};
      float dt = -1.0f;
#pragma unroll
      for (int i = 0; i < PF_MAX_TEXELS; i++) w1[i] = 0.0f;
      if (shade) {
        dpn = dot(pn, ray.d);
        cn = dot(co, pn);
        ts0 = acc_div(cn, dpn);
        if (heights) {
          const float tq0 = dpn >= 0.0f ? tn : sel_max(tn, ts0);
          const f3 x0 = ray.d * tq0 - co;
          float hsum = 0.0f, s0 = 0.0f;
#pragma unroll
          for (int i = 0; i < PF_MAX_TEXELS; i++) {
            const float4 ph = s_tex[i].pos_h;
            const f3 df{x0.x - ph.x, x0.y - ph.y, x0.z - ph.z};
            const float w = fast_exp(dot(df, df) * kexp);
            hsum += w * ph.w;
            s0 += w;
          }
          ws0 = sel_max(s0, 1e-20f);
          height = hsum * fast_rcp(ws0);
          ts1 = acc_div(cn + height, dpn);
        } else {
          ts1 = ts0;
        }
        const float tq = dpn >= 0.0f ? tn : sel_max(tn, ts1);
        const f3 x1 = ray.d * tq - co;
        f3 rs{0.0f, 0.0f, 0.0f};
        float s1 = 0.0f;
#pragma unroll
        for (int i = 0; i < PF_MAX_TEXELS; i++) {
          const float4 ph = s_tex[i].pos_h;
          const float4 cl = s_tex[i].rgb;
          const f3 df{x1.x - ph.x, x1.y - ph.y, x1.z - ph.z};
          const float w = fast_exp(dot(df, df) * kexp);
          w1[i] = w;
          rs = rs + f3{cl.x, cl.y, cl.z} * w;
          s1 += w;
        }
        ws = sel_max(s1, 1e-20f);
        if (!heights) ws0 = ws;
        col = rs * fast_rcp(ws);
        if (dpn >= 0.0f && ts1 < tf) {
          tf_id = n_adj;
          tf = ts1;
        }
        if (dpn < 0.0f && ts1 > tn) {
          tn_id = n_adj;
          tn = ts1;
        }
        dt = tf - tn;
      }
      const bool active = shade && dt > 0.0f;
      if (ballot(active) == 0) continue;
      if (DSUM) {
        // the dL/dlog T increments of the adjoint below, nothing else
        if (active) {
          const float dlt = -sigma * dt;
          log_t -= dlt;
          const float alpha = 1.0f - fast_exp(dlt);
          const float trans = fast_exp(log_t);
          const float ndv = dpn;
          const float ndt = exact ? ndv * ndv : ndv;
          float dLdtrans = dot(g_rgb, col) * alpha;
          if (ndv > 0.0f) dLdtrans += g_nd * ndt * alpha;
          dLdtrans += s_gc[kk] * alpha * p.inv_hw;
          dLdtrans += dot(g_n, pn) * alpha;
          dLdlog_t += dLdtrans * trans;
          while (qidx >= 0 && trans > qcur) {
            dLdlog_t += p.g_qdepth[px * p.Q + qidx] / (sigma + q_eps);
            qidx--;
            if (qidx >= 0) qcur = p.depth_q[px * p.Q + qidx];
          }
        }
        continue;
      }

      // ---- compositing adjoint (rasterize.py:1366-1430)
      float dLdtn = 0.0f, dLdtf = 0.0f, dLdr = 0.0f, dLdsig = 0.0f;
      f3 dLdc{0.0f, 0.0f, 0.0f}, dLdnorm{0.0f, 0.0f, 0.0f};
      // what the per-texel gradient pass needs (zero on inactive lanes)
      f3 dLdcol{0.0f, 0.0f, 0.0f}, x1{0.0f, 0.0f, 0.0f}, x0{0.0f, 0.0f, 0.0f};
      float gdp_in = 0.0f, gdp0_in = 0.0f, dLdheight = 0.0f, inv_ws = 0.0f, inv_ws0 = 0.0f;
      const float kw = 20.0f * fast_rcp(r * r);
      int fid0 = -1, fid1 = -1;
      float4 fg0 = make_float4(0, 0, 0, 0), fg1 = make_float4(0, 0, 0, 0);
      if (active) {
        const float dlt = -sigma * dt;
        log_t -= dlt;
        const float e_dlt = fast_exp(dlt);
        const float alpha = 1.0f - e_dlt;
        const float trans = fast_exp(log_t);
        const float ndv = dpn;
        float dLdlog_t_i = dLdlog_t;
        float dLddlt = dLdlog_t;
        const float gc = s_gc[kk];
        const float cdot = dot(g_rgb, col);
        const float ndt = exact ? ndv * ndv : ndv;
        const float ndot = dot(g_n, pn);
        float dLdalpha = cdot * trans;
        if (ndv > 0.0f) dLdalpha += g_nd * ndt * trans;
        dLdalpha += gc * trans * p.inv_hw;
        dLdalpha += ndot * trans;
        float dLdtrans = cdot * alpha;
        if (ndv > 0.0f) dLdtrans += g_nd * ndt * alpha;
        dLdtrans += gc * alpha * p.inv_hw;
        dLdtrans += ndot * alpha;
        dLdcol = g_rgb * (trans * alpha);
        dLdlog_t_i += dLdtrans * trans;
        dLddlt += -dLdalpha * e_dlt;
        dLdsig += dLddlt * -dt;
        dLdtn += dLddlt * sigma;
        dLdtf += dLddlt * -sigma;
        while (qidx >= 0 && trans > qcur) {
          const float g = p.g_qdepth[px * p.Q + qidx];
          dLdsig += -g * __logf(trans * fast_rcp(qcur)) / (sigma * sigma + q_eps);
          dLdtn += g;
          dLdlog_t_i += g / (sigma + q_eps);
          qidx--;
          if (qidx >= 0) qcur = p.depth_q[px * p.Q + qidx];
        }
        dLdnorm = dLdnorm + g_n * (alpha * trans);
        if (ndv > 0.0f) dLdnorm = dLdnorm + ray.d * (g_nd * alpha * trans * (exact ? 2.0f * ndv : 1.0f));
        dLdlog_t = dLdlog_t_i;

        // ---- plane_intersection_bwd (texture.py:74-186)
        const float dLdts_in = (tn_id == n_adj) ? dLdtn : ((tf_id == n_adj) ? dLdtf : 0.0f);
        inv_ws = fast_rcp(ws);
        inv_ws0 = fast_rcp(ws0);
        f3 pdc{0.0f, 0.0f, 0.0f}, pdn{0.0f, 0.0f, 0.0f};
        // colour stage: the query-point gradient (the per-texel values are
        // recomputed in the texel pass below)
        const float tq = dpn >= 0.0f ? tn0 : sel_max(tn0, ts1);
        x1 = ray.d * tq - co;
        f3 dLdx{0.0f, 0.0f, 0.0f};
        gdp_in = dot(dLdcol, col);
#pragma unroll
        for (int i = 0; i < PF_MAX_TEXELS; i++) {
          // a texel under the cut-off adds an exact zero (select, no branch)
          const float wn = w1[i] * inv_ws;
          const float4 ph = s_tex[i].pos_h;
          const float4 cl = s_tex[i].rgb;
          const f3 df{x1.x - ph.x, x1.y - ph.y, x1.z - ph.z};
          const float gdp = dLdcol.x * cl.x + dLdcol.y * cl.y + dLdcol.z * cl.z;
          const float sc = !(wn < cut) ? kw * wn * (gdp - gdp_in) : 0.0f;
          dLdx = dLdx - df * sc;
        }
        const float dLdtq = dot(dLdx, ray.d);
        float dtn_q = (dpn >= 0.0f || tn0 >= ts1) ? dLdtq : 0.0f;
        float dts = (dpn < 0.0f && tn0 < ts1) ? dLdtq : 0.0f;
        dts += dLdts_in;
        if (fabsf(dts) > 0.0f) {
          // ray_plane_intersect_bwd with p = c, h = height (rendering_math.py:138-147)
          const float iden = 1.0f / (dpn + pl_eps);
          pdc = pdc + pn * (dts * iden);
          pdn = pdn + (co * iden - ray.d * ((cn + height) / (dpn * dpn + pl_eps))) * dts;
          dLdheight += dts * iden;
        }
        // height stage
        const float tq0 = dpn >= 0.0f ? tn0 : sel_max(tn0, ts0);
        x0 = ray.d * tq0 - co;
        f3 dLdx0{0.0f, 0.0f, 0.0f};
        gdp0_in = dLdheight * height;
        // heights all zero: every term below is zero (ph.w = 0, height = 0)
        if (heights) {
#pragma unroll
          for (int i = 0; i < PF_MAX_TEXELS; i++) {
            const float4 ph = s_tex[i].pos_h;
            const f3 df{x0.x - ph.x, x0.y - ph.y, x0.z - ph.z};
            const float w = fast_exp(dot(df, df) * kexp);
            // the cut-off tests the colour-stage weight sum, as the reference
            const float sc = !(w * inv_ws < cut) ? kw * (w * inv_ws0) * (dLdheight * ph.w - gdp0_in) : 0.0f;
            dLdx0 = dLdx0 - df * sc;
          }
        }
        const float dLdtq0 = dot(dLdx0, ray.d);
        dtn_q += (dpn >= 0.0f || tn0 >= ts0) ? dLdtq0 : 0.0f;
        const float dts0 = (dpn < 0.0f && tn0 < ts0) ? dLdtq0 : 0.0f;
        if (fabsf(dts0) > 0.0f) {
          const float iden = 1.0f / (dpn + pl_eps);
          pdc = pdc + pn * (dts0 * iden);
          pdn = pdn + (co * iden - ray.d * (cn / (dpn * dpn + pl_eps))) * dts0;
        }
        dLdc = dLdc + pdc;
        dLdnorm = dLdnorm + pdn;
        if (tn_id != n_adj) dtn_q += dLdtn;
        dLdtn = dtn_q;

        // ---- route dL/dt_near and dL/dt_far to sphere / faces (rasterize.py:1464-1505)
#pragma unroll
        for (int side = 0; side < 2; side++) {
          const int tid = side == 0 ? tn0_id : tf0_id;
          const float dLdt = side == 0 ? dLdtn : dLdtf;
          if (exact && side == 1 && tf_id == n_adj) continue;
          if (tid == -1) {
            // ray_sphere_intersect_bwd (rendering_math.py:26-59)
            const float qa = dot(ray.d, ray.d);
            const float qb = 2.0f * dot(oc, ray.d);
            const f3 cr = cross(oc, ray.d);
            const float disc = 4.0f * (qa * r * r - dot(cr, cr));
            if (disc >= 0.0f) {
              const float sq = __builtin_sqrtf(disc);
              const float tfar = (-qb + sq) / (2.0f * qa);
              const float tnear = (-qb - sq) / (2.0f * qa);
              if (!(tnear < 0.0f && tfar < 0.0f)) {
                const f3 w = (oc * (2.0f * qa) - ray.d * qb) * (1.0f / (qa * sq));
                const f3 dq = ray.d * (1.0f / qa);
                if (side == 0) {
                  if (!(tnear < 0.0f)) {
                    dLdc = dLdc + (dq - w) * dLdt;
                    dLdr += dLdt * (-2.0f * r / sq);
                  }
                } else {
                  dLdc = dLdc + (dq + w) * dLdt;
                  dLdr += dLdt * (2.0f * r / sq);
                }
              }
            }
          } else if (tid >= 0 && tid < n_adj) {
            // ray_pface_intersect_bwd (rendering_math.py:101-125) with the
            // neighbour's full-precision sphere
            // the id staged with the face when the list fits one chunk
            const int adj = n_adj <= WAVE ? lds_load(s_fadj + tid) : p.adjacency[sr.a0 + tid];
            const float4 as = p.spheres[adj];
            const f3 a{as.x, as.y, as.z};
            const f3 n = a - c;
            const float num = dot(co, n) + 0.5f * (dot(n, n) + r * r - as.w * as.w);
            const float idp = 1.0f / dot(ray.d, n);
            const f3 qv = ray.d * (num * idp * idp);
            dLdc = dLdc + (oc * idp + qv) * dLdt;
            dLdr += dLdt * (r * idp);
            const f3 ga = ((a - ray.o) * idp - qv) * dLdt;
            const float4 g4 = make_float4(ga.x, ga.y, ga.z, dLdt * (-as.w * idp));
            if (side == 0) {
              fid0 = tid;
              fg0 = g4;
            } else {
              fid1 = tid;
              fg1 = g4;
            }
          }
        }
      }

      // ---- neighbour gradients, summed over the wave per face in
      // registers: the rays of a tile mostly leave through the same few
      // faces, so the distinct (side, face id) pairs of the entry and exit
      // faces are taken four at a time (ballot), each lane's four values go
      // to its face's slots of a 16-value slice and wave_sum16 sums them;
      // one atomic per face value. (Per-lane LDS atomics into a slot per
      // face serialise on those few addresses.)
      {
        uint64_t pend0 = ballot(fid0 >= 0), pend1 = ballot(fid1 >= 0);
        while (pend0 | pend1) {
          int fk[4];
          float v[RED_SLICE];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            int f = -1;
            bool in = false;
            float4 g = fg0;
            if (pend0) {
              f = __builtin_amdgcn_readlane(fid0, __builtin_ctzll(pend0));
              in = fid0 == f;
              pend0 &= ~ballot(in);
            } else if (pend1) {
              f = __builtin_amdgcn_readlane(fid1, __builtin_ctzll(pend1));
              in = fid1 == f;
              pend1 &= ~ballot(in);
              g = fg1;
            }
            fk[k] = f;
            v[4 * k + 0] = in ? g.x : 0.0f;
            v[4 * k + 1] = in ? g.y : 0.0f;
            v[4 * k + 2] = in ? g.z : 0.0f;
            v[4 * k + 3] = in ? g.w : 0.0f;
          }
          const float acc = wave_sum16(v, lane);
          // lane 4u holds value u: component u % 4 of face slot u / 4
          const int u = lane >> 2, k = u >> 2;
          const int f = k == 0 ? fk[0] : (k == 1 ? fk[1] : (k == 2 ? fk[2] : fk[3]));
          if ((lane & 3) == 0 && f >= 0 && acc != 0.0f) {
            const int adj = n_adj <= WAVE ? lds_load(s_fadj + f) : p.adjacency[sr.a0 + f];
            atomicAdd(p.g_spheres + 4 * (size_t)adj + (u & 3), acc);
          }
        }
      }

      // ---- texel and site gradients: summed over the wave in registers
      // (wave_sum16) in four 16-value slices, one atomic per value: slice j
      // holds the 14 values of texels 2j and 2j + 1 (slots 0-13) and the
      // site's own values 2j and 2j + 1 (slots 14-15), so the 8 + 7 * 8
      // values of a site fill four reductions. Texel values are produced two
      // texels per slice, so no per-texel accumulators stay live across the
      // adjoint.
      const float site_v[8] = {dLdc.x, dLdc.y, dLdc.z, dLdr, dLdnorm.x, dLdnorm.y, dLdnorm.z, dLdsig};
      // Branch-free: a skipped texel (w / sum w < cut, or an inactive lane)
      // selects zero, so no per-texel control flow forms in this pass.
#pragma unroll
      for (int i0 = 0; i0 < PF_MAX_TEXELS; i0 += 2) {
        {
          float v[RED_SLICE];
          v[14] = site_v[i0];
          v[15] = site_v[i0 + 1];
#pragma unroll
          for (int k = 0; k < 2; k++) {
            // i <= 7 stays inside s_tex; values past T are never flushed
            const int i = i0 + k;
            const float4 ph = s_tex[i].pos_h;
            const float4 cl = s_tex[i].rgb;
            // colour stage (texture.py:158-171)
            const float wn = w1[i] * inv_ws;
            const bool use1 = active && !(wn < cut);
            const f3 df{x1.x - ph.x, x1.y - ph.y, x1.z - ph.z};
            const float gdp = dLdcol.x * cl.x + dLdcol.y * cl.y + dLdcol.z * cl.z;
            const float s1 = kw * wn * (gdp - gdp_in);
            float sx = use1 ? df.x * s1 : 0.0f, sy = use1 ? df.y * s1 : 0.0f, sz = use1 ? df.z * s1 : 0.0f;
            v[7 * k + 3] = use1 ? dLdcol.x * wn : 0.0f;
            v[7 * k + 4] = use1 ? dLdcol.y * wn : 0.0f;
            v[7 * k + 5] = use1 ? dLdcol.z * wn : 0.0f;
            // height stage (texture.py:118-131); with all heights zero it
            // queries the colour-stage point and its site terms vanish
            if (heights) {
              const f3 d0{x0.x - ph.x, x0.y - ph.y, x0.z - ph.z};
              const float w = fast_exp(dot(d0, d0) * kexp);
              const bool use0 = active && !(w * inv_ws < cut);
              const float wn0 = w * inv_ws0;
              const float s0 = kw * wn0 * (dLdheight * ph.w - gdp0_in);
              v[7 * k + 6] = use0 ? wn0 * dLdheight : 0.0f;
              sx += use0 ? d0.x * s0 : 0.0f;
              sy += use0 ? d0.y * s0 : 0.0f;
              sz += use0 ? d0.z * s0 : 0.0f;
            } else {
              v[7 * k + 6] = use1 ? (w1[i] * inv_ws0) * dLdheight : 0.0f;
            }
            v[7 * k + 0] = sx;
            v[7 * k + 1] = sy;
            v[7 * k + 2] = sz;
          }
          // texels past T hold zeros and are not flushed
#pragma unroll
          for (int q = 0; q < 14; q++)
            if (i0 + q / 7 >= T) v[q] = 0.0f;
          const float acc = wave_sum16(v, lane);
          const int u = lane >> 2;  // this lane's value of the slice
          const bool valid = u >= 14 || i0 + u / 7 < T;
          const int slot = u >= 14 ? i0 + (u - 14) : 8 + 7 * i0 + u;
          if ((lane & 3) == 0 && valid && acc != 0.0f) atomicAdd(grad_slot(p, sr.site, T, slot), acc);
        }
      }
    }
  }
  if (DSUM) p.seg_dsum[(size_t)wi.seg * WAVE + lane] = dLdlog_t;
}


// ------------------------------------------------------ benchmark forward
// benchmark_kernel (rasterize.py:614-734): colour only; a lane stops once its
// T < threshold before a site, the wave once no lane is left; faces come from
// the (c_adj - c, pm_adj - pm) encoding (ray_pface_intersect_diff,
// rendering_math.py:85-98) and a lane's face loop ends with its interval.

struct RgbParams {
  RasterCam cam;
  int n_texels, n_tiles;
  const float4* spheres;
  const float4* nsigmas;
  const float* texel_sites;
  const float4* texel_rgbh;
  const uint2* texel_rgbh_bf16;
  const int32_t* adjacency;
  const int32_t* adj_off;
  const uint16_t* adj_diff;
  const int64_t* offsets;
  const int32_t* sites;
  const int32_t* order;
  float thr;
  float* color;
  // visualisation only
  int max_inter;
  float depth_q;
  f3 bkgd;
  float* depth;
  float* normal;
  float* alpha;
  int32_t* inter;
};

// Colour-only compositing with a per-ray stop. VIS = false: the benchmark
// kernel (rasterize.py:614-734), faces from the (c_adj - c, pm_adj - pm)
// encoding. VIS = true: the visualisation kernel (rasterize.py:739-890), faces
// from the training encoding (c_adj - c, r_adj), plus background, depth at
// T = depth_q, normals, alpha and per-ray entry counts.
template <bool PINHOLE, bool VIS>
__global__ void __launch_bounds__(WAVE* WPB) raster_rgb_kernel(RgbParams p) {
  __shared__ SiteRec s_site_all[WPB][WAVE];
  __shared__ float4 s_face_all[WPB][WAVE];
  __shared__ TexelLds s_tex_all[WPB][PF_MAX_TEXELS];

  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int tile = wave_tile(p.order, p.n_tiles, blockIdx.x, wv);
  if (tile < 0) return;
  SiteRec* s_site = s_site_all[wv];
  float4* s_face = s_face_all[wv];
  TexelLds* s_tex = s_tex_all[wv];

  const Ray ray = lane_ray<PINHOLE>(p.cam, tile, lane);
  const int T = p.n_texels;
  float log_t = 0.0f;
  f3 rgb{0.0f, 0.0f, 0.0f};
  f3 nacc{0.0f, 0.0f, 0.0f};
  float depth = 0.0f;
  bool found = false;
  int nhit = 0;
  bool alive = !ray.oob;  // out-of-image lanes are never written

  const int64_t start = p.offsets[tile];
  const int total = (int)(p.offsets[tile + 1] - start);
  for (int base = 0; base < total && ballot(alive) != 0; base += WAVE) {
    const int cnt = min(WAVE, total - base);
    wave_sync();
    if (lane < cnt) {
      const int s = p.sites[start + base + lane];
      s_site[lane].sph = p.spheres[s];
      s_site[lane].nsg = p.nsigmas[s];
      s_site[lane].site = s;
      s_site[lane].a0 = p.adj_off[s];
      s_site[lane].a1 = p.adj_off[s + 1];
    }
    wave_sync();
    for (int kk = 0; kk < cnt; kk++) {
      const float trans = fast_exp(log_t);
      if (VIS) {
        // the entry that stops a ray is counted (rasterize.py:790-797)
        nhit += alive ? 1 : 0;
        alive = alive && base + kk < p.max_inter;
      }
      alive = alive && trans >= p.thr;
      if (ballot(alive) == 0) break;
      const SiteRec sr = uniform_site(s_site[kk]);
      const f3 c{sr.sph.x, sr.sph.y, sr.sph.z};
      const float r = sr.sph.w;
      const f3 pn{sr.nsg.x, sr.nsg.y, sr.nsg.z};
      const float sigma = sr.nsg.w;
      const f3 oc = ray.o - c;
      const f3 co = -oc;
      float tn, tf;
      bool open = alive && sphere_hit(oc, ray.d, r, tn, tf);
      if (ballot(open) == 0) continue;
      for (int e0 = sr.a0; e0 < sr.a1 && ballot(open) != 0; e0 += WAVE) {
        const int ec = __builtin_amdgcn_readfirstlane(min(WAVE, sr.a1 - e0));
        const int ec4 = (ec + 3) & ~3;
        wave_sync();
        if (lane < ec) {
          if (VIS) {
            s_face[lane] = face_of<PINHOLE>(p.adjacency, p.adj_diff, p.spheres, e0 + lane, sr.sph, c - p.cam.eye);
          } else {
            const uint2 raw = *reinterpret_cast<const uint2*>(p.adj_diff + 4 * (int64_t)(e0 + lane));
            const f3 n{half_bits_to_float((uint16_t)(raw.x & 0xffff)), half_bits_to_float((uint16_t)(raw.x >> 16)),
                       half_bits_to_float((uint16_t)(raw.y & 0xffff))};
            float w = half_bits_to_float((uint16_t)(raw.y >> 16));
            if (PINHOLE) w -= dot(p.cam.eye, n);
            s_face[lane] = make_float4(n.x, n.y, n.z, w);
          }
        } else if (lane < ec4) {
          s_face[lane] = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));  // no ray crosses it (as stage_faces)
        }
        wave_sync();
        if (open) {
          // four faces per LDS round trip; a lane leaves once its interval is
          // empty (later faces only narrow an empty interval further)
          for (int e = 0; e < ec4; e += 4) {
            float4 fq[4];
#pragma unroll
            for (int j = 0; j < 4; j++) fq[j] = s_face[e + j];
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const float4 f = fq[j];
              const f3 n{f.x, f.y, f.z};
              const float dp = dot(ray.d, n);
              const float num = PINHOLE ? f.w : (VIS ? f.w + dot(co, n) : f.w - dot(ray.o, n));
              const float t = acc_div(num, dp);
              tf = (dp >= 0.0f && t < tf) ? t : tf;
              tn = (dp < 0.0f && t > tn) ? t : tn;
            }
            if (tn > tf) break;
          }
          open = !(tn > tf);
        }
      }
      if (ballot(open) == 0) continue;
      wave_sync();
      const bool heights = stage_texels(s_tex, p.texel_sites, p.texel_rgbh, p.texel_rgbh_bf16, sr.site, T, c, lane);
      wave_sync();
      if (open) {
        // plane_intersection_fwd_local (rasterize.py:67-138)
        const float dpn = dot(pn, ray.d);
        const float cn = dot(co, pn);
        const float kexp = -10.0f * fast_rcp(r * r);
        const float ts0 = acc_div(cn, dpn);
        float height = 0.0f;
        if (heights) {
          const float tq0 = dpn >= 0.0f ? tn : sel_max(tn, ts0);
          const f3 x0 = ray.d * tq0 - co;
          float hsum = 0.0f, ws0 = 0.0f;
#pragma unroll
          for (int i = 0; i < PF_MAX_TEXELS; i++) {
            const float4 ph = s_tex[i].pos_h;
            const f3 df{x0.x - ph.x, x0.y - ph.y, x0.z - ph.z};
            const float w = fast_exp(dot(df, df) * kexp);
            hsum += w * ph.w;
            ws0 += w;
          }
          height = hsum * fast_rcp(sel_max(ws0, 1e-20f));
        }
        const float ts1 = heights ? acc_div(cn + height, dpn) : ts0;
        const float tq = dpn >= 0.0f ? tn : sel_max(tn, ts1);
        const f3 x1 = ray.d * tq - co;
        f3 rs{0.0f, 0.0f, 0.0f};
        float ws = 0.0f;
#pragma unroll
        for (int i = 0; i < PF_MAX_TEXELS; i++) {
          const float4 ph = s_tex[i].pos_h;
          const float4 cl = s_tex[i].rgb;
          const f3 df{x1.x - ph.x, x1.y - ph.y, x1.z - ph.z};
          const float w = fast_exp(dot(df, df) * kexp);
          rs = rs + f3{cl.x, cl.y, cl.z} * w;
          ws += w;
        }
        const f3 col = rs * fast_rcp(sel_max(ws, 1e-20f));
        const float tf2 = (dpn >= 0.0f && ts1 < tf) ? ts1 : tf;
        const float tn2 = (dpn < 0.0f && ts1 > tn) ? ts1 : tn;
        const float dt = tf2 - tn2;
        if (dt > 0.0f) {
          const float dlt = -sigma * dt;
          const float at = (1.0f - fast_exp(dlt)) * trans;
          rgb = rgb + col * at;
          log_t += dlt;
          if (VIS) {
            nacc = nacc + pn * at;
            if (!found && fast_exp(log_t) < p.depth_q) {
              depth = tn2 + logf(trans / p.depth_q) / sigma;
              found = true;
            }
          }
        }
      }
    }
  }
  if (!ray.oob) {
    const size_t px = (size_t)ray.pi * p.cam.W + ray.pj;
    if (VIS) {
      const float tr = fast_exp(log_t);
      rgb = rgb + p.bkgd * tr;
      p.depth[px] = depth;
      p.normal[3 * px + 0] = nacc.x;
      p.normal[3 * px + 1] = nacc.y;
      p.normal[3 * px + 2] = nacc.z;
      p.alpha[px] = 1.0f - tr;
      p.inter[px] = nhit;
    }
    p.color[3 * px + 0] = rgb.x;
    p.color[3 * px + 1] = rgb.y;
    p.color[3 * px + 2] = rgb.z;
  }
}

__global__ void __launch_bounds__(256) scrub_kernel(float* x, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !isfinite(x[i])) x[i] = 0.0f;
}

// The same over float4 words (a 16-byte aligned buffer of n4 words): a read
// of every value and a write only where one is not finite.
__global__ void __launch_bounds__(256) scrub4_kernel(float4* x, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    float4 v = x[i];
    if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w))) {
      v.x = isfinite(v.x) ? v.x : 0.0f;
      v.y = isfinite(v.y) ? v.y : 0.0f;
      v.z = isfinite(v.z) ? v.z : 0.0f;
      v.w = isfinite(v.w) ? v.w : 0.0f;
      x[i] = v;
    }
  }
}

// The per-site words the compositing kernels read, from the separate
// parameter arrays: sphere (x, y, z, r), normal and density (nx, ny, nz, σ)
// and per texel (r, g, b, height). One thread per texel row; the first
// n_sites threads also write their site's two words.
__global__ void __launch_bounds__(256)
    pack_sites_kernel(const float* __restrict__ points, const float* __restrict__ radii,
                      const float* __restrict__ normals, const float* __restrict__ density,
                      const float* __restrict__ rgb, const float* __restrict__ height, int64_t n, int64_t rows,
                      float4* __restrict__ spheres, float4* __restrict__ nsigmas, float4* __restrict__ rgbh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows)
    rgbh[i] = rgb ? make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], height[i])
                  : make_float4(0.0f, 0.0f, 0.0f, height[i]);
  if (i < n) {
    spheres[i] = make_float4(points[3 * i], points[3 * i + 1], points[3 * i + 2], radii[i]);
    nsigmas[i] = make_float4(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2], density[i]);
  }
}

}  // namespace pf

using namespace pf;

static int check_scene(const pf_scene& s, const char* who) {
  if (s.n_sites < 0 || (s.n_sites > 0 && (!s.spheres || !s.nsigmas || !s.texel_sites ||
                                          (!s.texel_rgbh && !s.texel_rgbh_bf16) || !s.adjacency_offsets)))
    return set_error(PF_ERR_ARG, "%s: missing scene arrays", who);
  if (reinterpret_cast<uintptr_t>(s.texel_rgbh_bf16) % 8)
    return set_error(PF_ERR_ARG, "%s: texel_rgbh_bf16 must be 8-byte aligned", who);
  if (s.n_texels < 1 || s.n_texels > PF_MAX_TEXELS)
    return set_error(PF_ERR_ARG, "%s: n_texels=%d outside 1..%d", who, s.n_texels, PF_MAX_TEXELS);
  return PF_OK;
}

extern "C" {

int pf_raster_fwd(const pf_raster_fwd_args* a, void* stream) {
  if (!a) return set_error(PF_ERR_ARG, "pf_raster_fwd: null args");
  int st = check_scene(a->scene, "pf_raster_fwd");
  if (st) return st;
  const pf_camera& c = a->cam;
  if (c.width < 2 || c.height < 2) return set_error(PF_ERR_ARG, "pf_raster_fwd: image must be >= 2x2");
  const int nt = tiles_h(c) * tiles_w(c);
  if (a->tiles.n_tiles != nt || !a->tiles.offsets)
    return set_error(PF_ERR_ARG, "pf_raster_fwd: tile lists do not match the camera (%d vs %d)", a->tiles.n_tiles, nt);
  // per-site arrays may be null for a scene without sites (no list entry reads them)
  const bool sites = a->scene.n_sites > 0;
  if (!a->tile_early_stop || !a->color || !a->log_t || !a->normal_distance || !a->normal || (sites && !a->contrib))
    return set_error(PF_ERR_ARG, "pf_raster_fwd: missing outputs");
  if ((a->return_point_err || a->loss_accumulation) && !a->ray_gt)
    return set_error(PF_ERR_ARG, "pf_raster_fwd: ray_gt is required for point error / surface loss");
  if (a->return_point_err && sites && !a->point_err) return set_error(PF_ERR_ARG, "pf_raster_fwd: missing point_err");
  if (a->depth_quantiles && (a->n_quantiles < 1 || !a->quantile_depths))
    return set_error(PF_ERR_ARG, "pf_raster_fwd: quantiles need n_quantiles >= 1 and an output");
  FwdParams p{};
  p.cam = make_rcam(c);
  p.n_texels = a->scene.n_texels;
  p.n_tiles = nt;
  p.Q = a->depth_quantiles ? a->n_quantiles : 0;
  p.spheres = reinterpret_cast<const float4*>(a->scene.spheres);
  p.nsigmas = reinterpret_cast<const float4*>(a->scene.nsigmas);
  p.texel_sites = a->scene.texel_sites;
  p.texel_rgbh = reinterpret_cast<const float4*>(a->scene.texel_rgbh);
  p.texel_rgbh_bf16 = reinterpret_cast<const uint2*>(a->scene.texel_rgbh_bf16);
  p.adjacency = a->scene.adjacency;
  p.adj_off = a->scene.adjacency_offsets;
  p.adj_diff = a->scene.adjacency_diff;
  p.adj_cnt = a->scene.adjacency_count;
  p.offsets = a->tiles.offsets;
  p.sites = a->tiles.sites;
  p.order = a->tiles.order;
  p.depth_q = a->depth_quantiles;
  p.ray_gt = a->ray_gt;
  p.thr = a->transmittance_threshold;
  p.inv_hw = 1.0f / (float)(c.height * c.width);
  p.point_err_on = a->return_point_err;
  p.loss_acc = a->loss_accumulation;
  p.early = a->tile_early_stop;
  p.color = a->color;
  p.log_t = a->log_t;
  p.ndist = a->normal_distance;
  p.normal = a->normal;
  p.qdepth = a->quantile_depths;
  p.err = a->err;
  p.contrib = a->contrib;
  p.point_err = a->point_err;
  p.entry_shades = a->entry_shades;
  const pf_mlp_head* hd = a->mlp_head;
  if (hd) {
    const pf_mlp_weights& w = hd->weights;
    if (!w.w1 || !w.b1 || !w.w2 || !w.b2 || !w.w3 || !w.b3 || (sites && (!hd->latents || !hd->texel_rgbh_out)))
      return set_error(PF_ERR_ARG, "pf_raster_fwd: the learned head needs its weights, latents and texel_rgbh_out");
    if ((reinterpret_cast<uintptr_t>(hd->latents) | reinterpret_cast<uintptr_t>(hd->texel_rgbh_out)) % 16)
      return set_error(PF_ERR_ARG, "pf_raster_fwd: latents and texel_rgbh_out must be 16-byte aligned");
    if (hd->encoding != PF_MLP_ENC_SH && hd->encoding != PF_MLP_ENC_FREQ)
      return set_error(PF_ERR_ARG, "pf_raster_fwd: unknown direction encoding %d", hd->encoding);
    p.mlp_w1 = w.w1;
    p.mlp_w2 = w.w2;
    p.mlp_w3 = w.w3;
    p.mlp_b1 = w.b1;
    p.mlp_b2 = w.b2;
    p.mlp_b3 = w.b3;
    p.mlp_lat = hd->latents;
    p.mlp_enc = hd->encoding;
    p.rgbh_out = reinterpret_cast<float4*>(hd->texel_rgbh_out);
  }
  const pf_segments& sg = a->segs;
  hipStream_t hs = as_stream(stream);
  if (sg.n_segs > 0) {
    if (sg.seg_len < WAVE || sg.seg_len % WAVE || !sg.tile_first || !sg.seg_tile || !sg.order || !sg.log_t_sums ||
        !sg.lane_active || !sg.scratch || sg.n_segs < nt)
      return set_error(PF_ERR_ARG, "pf_raster_fwd: bad segment table");
    p.n_segs = sg.n_segs;
    p.seg_count = sg.count;
    p.seg_len = sg.seg_len;
    p.seg_first = sg.tile_first;
    p.seg_tile = sg.seg_tile;
    p.seg_order = sg.order;
    p.seg_logt = sg.log_t_sums;
    p.seg_final = sg.scratch;
    p.seg_active = sg.lane_active;
    fill_int_kernel<<<(nt + 255) / 256, 256, 0, hs>>>(a->tile_early_stop, nt, 0x7fffffff);
    const int grid = (sg.n_segs + WPB - 1) / WPB;
    // the log-T prepass reads heights only: no head there
    const int hgrid = (sg.n_segs + HEAD_WPB - 1) / HEAD_WPB;
    if (c.ray_maps) {
      raster_fwd_kernel<false, true><<<grid, WAVE * WPB, 0, hs>>>(p);
      if (hd)
        raster_fwd_kernel<false, false, true><<<hgrid, WAVE * HEAD_WPB, 0, hs>>>(p);
      else
        raster_fwd_kernel<false, false><<<grid, WAVE * WPB, 0, hs>>>(p);
      raster_fwd_finish_kernel<false><<<nt, WAVE, 0, hs>>>(p);
    } else {
      raster_fwd_kernel<true, true><<<grid, WAVE * WPB, 0, hs>>>(p);
      if (hd)
        raster_fwd_kernel<true, false, true><<<hgrid, WAVE * HEAD_WPB, 0, hs>>>(p);
      else
        raster_fwd_kernel<true, false><<<grid, WAVE * WPB, 0, hs>>>(p);
      raster_fwd_finish_kernel<true><<<nt, WAVE, 0, hs>>>(p);
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
  }
  if (hd) {
    const int grid = xcd_grid(nt, HEAD_WPB);
    if (c.ray_maps)
      raster_fwd_kernel<false, false, true><<<grid, WAVE * HEAD_WPB, 0, hs>>>(p);
    else
      raster_fwd_kernel<true, false, true><<<grid, WAVE * HEAD_WPB, 0, hs>>>(p);
  } else {
    const int grid = xcd_grid(nt, WPB);
    if (c.ray_maps)
      raster_fwd_kernel<false, false><<<grid, WAVE * WPB, 0, hs>>>(p);
    else
      raster_fwd_kernel<true, false><<<grid, WAVE * WPB, 0, hs>>>(p);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_raster_bwd(const pf_raster_bwd_args* a, void* stream) {
  if (!a) return set_error(PF_ERR_ARG, "pf_raster_bwd: null args");
  int st = check_scene(a->scene, "pf_raster_bwd");
  if (st) return st;
  const pf_camera& c = a->cam;
  const int nt = tiles_h(c) * tiles_w(c);
  if (a->tiles.n_tiles != nt || !a->tiles.offsets || !a->tile_early_stop)
    return set_error(PF_ERR_ARG, "pf_raster_bwd: tile lists do not match the camera");
  // per-site arrays may be null for a scene without sites (no list entry reads them)
  const bool sites = a->scene.n_sites > 0;
  if (!a->log_t || !a->grad_color || !a->grad_log_t || !a->grad_normal_distance || !a->grad_normal ||
      (sites && !a->grad_contrib))
    return set_error(PF_ERR_ARG, "pf_raster_bwd: missing incoming gradients");
  if (sites && (!a->grad_spheres || !a->grad_nsigmas || !a->grad_texel_sites || !a->grad_texel_rgb ||
                !a->grad_texel_height))
    return set_error(PF_ERR_ARG, "pf_raster_bwd: missing gradient outputs");
  if (a->depth_quantiles && (a->n_quantiles < 1 || !a->grad_quantile_depths))
    return set_error(PF_ERR_ARG, "pf_raster_bwd: quantiles need their gradient");
  BwdParams p{};
  p.cam = make_rcam(c);
  p.n_texels = a->scene.n_texels;
  p.n_tiles = nt;
  p.Q = a->depth_quantiles ? a->n_quantiles : 0;
  p.exact = a->exact_gradients;
  p.spheres = reinterpret_cast<const float4*>(a->scene.spheres);
  p.nsigmas = reinterpret_cast<const float4*>(a->scene.nsigmas);
  p.texel_sites = a->scene.texel_sites;
  p.texel_rgbh = reinterpret_cast<const float4*>(a->scene.texel_rgbh);
  p.texel_rgbh_bf16 = reinterpret_cast<const uint2*>(a->scene.texel_rgbh_bf16);
  p.adjacency = a->scene.adjacency;
  p.adj_off = a->scene.adjacency_offsets;
  p.adj_diff = a->scene.adjacency_diff;
  p.adj_cnt = a->scene.adjacency_count;
  p.offsets = a->tiles.offsets;
  p.sites = a->tiles.sites;
  p.order = a->tiles.order;
  p.depth_q = a->depth_quantiles;
  p.early = a->tile_early_stop;
  p.log_t = a->log_t;
  p.g_color = a->grad_color;
  p.g_log_t = a->grad_log_t;
  p.g_ndist = a->grad_normal_distance;
  p.g_normal = a->grad_normal;
  p.g_qdepth = a->grad_quantile_depths;
  p.g_contrib = a->grad_contrib;
  p.inv_hw = 1.0f / (float)(c.height * c.width);
  p.g_spheres = a->grad_spheres;
  p.g_nsigmas = a->grad_nsigmas;
  p.g_tsites = a->grad_texel_sites;
  p.g_trgb = a->grad_texel_rgb;
  p.g_theight = a->grad_texel_height;
  p.entry_shades = a->entry_shades;
  const pf_segments& sg = a->segs;
  hipStream_t hs = as_stream(stream);
  if (sg.n_segs > 0) {
    if (sg.seg_len < WAVE || sg.seg_len % WAVE || !sg.tile_first || !sg.seg_tile || !sg.order || !sg.log_t_sums ||
        !sg.lane_active || !sg.scratch || sg.n_segs < nt)
      return set_error(PF_ERR_ARG, "pf_raster_bwd: bad segment table");
    p.n_segs = sg.n_segs;
    p.seg_count = sg.count;
    p.seg_len = sg.seg_len;
    p.seg_first = sg.tile_first;
    p.seg_tile = sg.seg_tile;
    p.seg_order = sg.order;
    p.seg_logt = sg.log_t_sums;
    p.seg_active = sg.lane_active;
    p.seg_dsum = sg.scratch;
    const int grid = (sg.n_segs + WPB - 1) / WPB;
    if (c.ray_maps) {
      raster_bwd_kernel<false, true><<<grid, WAVE * WPB, 0, hs>>>(p);
      raster_bwd_kernel<false, false><<<grid, WAVE * WPB, 0, hs>>>(p);
    } else {
      raster_bwd_kernel<true, true><<<grid, WAVE * WPB, 0, hs>>>(p);
      raster_bwd_kernel<true, false><<<grid, WAVE * WPB, 0, hs>>>(p);
    }
    PF_HIP(hipGetLastError());
    return PF_OK;
  }
  const int groups = (nt + WPB - 1) / WPB;
  const int grid = ((groups + 7) / 8) * 8;
  if (c.ray_maps)
    raster_bwd_kernel<false, false><<<grid, WAVE * WPB, 0, hs>>>(p);
  else
    raster_bwd_kernel<true, false><<<grid, WAVE * WPB, 0, hs>>>(p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_raster_rgb(const pf_raster_rgb_args* a, void* stream) {
  if (!a) return set_error(PF_ERR_ARG, "pf_raster_rgb: null args");
  int st = check_scene(a->scene, "pf_raster_rgb");
  if (st) return st;
  const pf_camera& c = a->cam;
  if (c.width < 2 || c.height < 2) return set_error(PF_ERR_ARG, "pf_raster_rgb: image must be >= 2x2");
  const int nt = tiles_h(c) * tiles_w(c);
  if (a->tiles.n_tiles != nt || !a->tiles.offsets) return set_error(PF_ERR_ARG, "pf_raster_rgb: tile lists do not match the camera");
  if (!a->color) return set_error(PF_ERR_ARG, "pf_raster_rgb: missing colour output");
  if (a->scene.n_sites > 0 && a->scene.adjacency_offsets && !a->scene.adjacency_diff)
    return set_error(PF_ERR_ARG, "pf_raster_rgb: the power-face encoding is required");
  RgbParams p{};
  p.cam = make_rcam(c);
  p.n_texels = a->scene.n_texels;
  p.n_tiles = nt;
  p.spheres = reinterpret_cast<const float4*>(a->scene.spheres);
  p.nsigmas = reinterpret_cast<const float4*>(a->scene.nsigmas);
  p.texel_sites = a->scene.texel_sites;
  p.texel_rgbh = reinterpret_cast<const float4*>(a->scene.texel_rgbh);
  p.texel_rgbh_bf16 = reinterpret_cast<const uint2*>(a->scene.texel_rgbh_bf16);
  p.adj_off = a->scene.adjacency_offsets;
  p.adj_diff = a->scene.adjacency_diff;
  p.offsets = a->tiles.offsets;
  p.sites = a->tiles.sites;
  p.order = a->tiles.order;
  p.thr = a->transmittance_threshold;
  p.color = a->color;
  const int groups = (nt + WPB - 1) / WPB;
  const int grid = ((groups + 7) / 8) * 8;
  if (c.ray_maps)
    raster_rgb_kernel<false, false><<<grid, WAVE * WPB, 0, as_stream(stream)>>>(p);
  else
    raster_rgb_kernel<true, false><<<grid, WAVE * WPB, 0, as_stream(stream)>>>(p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_raster_vis(const pf_raster_vis_args* a, void* stream) {
  if (!a) return set_error(PF_ERR_ARG, "pf_raster_vis: null args");
  int st = check_scene(a->scene, "pf_raster_vis");
  if (st) return st;
  const pf_camera& c = a->cam;
  if (c.width < 2 || c.height < 2) return set_error(PF_ERR_ARG, "pf_raster_vis: image must be >= 2x2");
  const int nt = tiles_h(c) * tiles_w(c);
  if (a->tiles.n_tiles != nt || !a->tiles.offsets)
    return set_error(PF_ERR_ARG, "pf_raster_vis: tile lists do not match the camera");
  if (!a->color || !a->depth || !a->normal || !a->alpha || !a->intersections)
    return set_error(PF_ERR_ARG, "pf_raster_vis: missing outputs");
  if (a->scene.n_sites > 0 && a->scene.adjacency_offsets && !a->scene.adjacency_diff && !a->scene.adjacency)
    return set_error(PF_ERR_ARG, "pf_raster_vis: needs the adjacency encoding or the adjacency");
  RgbParams p{};
  p.cam = make_rcam(c);
  p.n_texels = a->scene.n_texels;
  p.n_tiles = nt;
  p.spheres = reinterpret_cast<const float4*>(a->scene.spheres);
  p.nsigmas = reinterpret_cast<const float4*>(a->scene.nsigmas);
  p.texel_sites = a->scene.texel_sites;
  p.texel_rgbh = reinterpret_cast<const float4*>(a->scene.texel_rgbh);
  p.texel_rgbh_bf16 = reinterpret_cast<const uint2*>(a->scene.texel_rgbh_bf16);
  p.adjacency = a->scene.adjacency;
  p.adj_off = a->scene.adjacency_offsets;
  p.adj_diff = a->scene.adjacency_diff;
  p.offsets = a->tiles.offsets;
  p.sites = a->tiles.sites;
  p.order = a->tiles.order;
  p.thr = a->transmittance_threshold;
  p.color = a->color;
  p.max_inter = a->max_intersections;
  p.depth_q = a->depth_quantile;
  p.bkgd = f3{a->bkgd_color[0], a->bkgd_color[1], a->bkgd_color[2]};
  p.depth = a->depth;
  p.normal = a->normal;
  p.alpha = a->alpha;
  p.inter = a->intersections;
  const int groups = (nt + WPB - 1) / WPB;
  const int grid = ((groups + 7) / 8) * 8;
  if (c.ray_maps)
    raster_rgb_kernel<false, true><<<grid, WAVE * WPB, 0, as_stream(stream)>>>(p);
  else
    raster_rgb_kernel<true, true><<<grid, WAVE * WPB, 0, as_stream(stream)>>>(p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_pack_sites(const float* points, const float* radii, const float* normals, const float* density,
                  const float* texel_rgb, const float* texel_height, int32_t n_sites, int32_t n_texels,
                  float* spheres, float* nsigmas, float* texel_rgbh, void* stream) {
  if (n_sites < 0 || n_texels < 1 || n_texels > PF_MAX_TEXELS ||
      (n_sites > 0 && (!points || !radii || !normals || !density || !texel_height || !spheres || !nsigmas ||
                       !texel_rgbh)))
    return set_error(PF_ERR_ARG, "pf_pack_sites: bad arguments");
  if ((reinterpret_cast<uintptr_t>(spheres) | reinterpret_cast<uintptr_t>(nsigmas) |
       reinterpret_cast<uintptr_t>(texel_rgbh)) % 16)
    return set_error(PF_ERR_ARG, "pf_pack_sites: outputs must be 16-byte aligned");
  if (n_sites == 0) return PF_OK;
  const int64_t rows = (int64_t)n_sites * n_texels;
  pack_sites_kernel<<<(unsigned)((rows + 255) / 256), 256, 0, as_stream(stream)>>>(
      points, radii, normals, density, texel_rgb, texel_height, n_sites, rows, reinterpret_cast<float4*>(spheres),
      reinterpret_cast<float4*>(nsigmas), reinterpret_cast<float4*>(texel_rgbh));
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_scrub_nonfinite(float* data, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && !data)) return set_error(PF_ERR_ARG, "pf_scrub_nonfinite: bad arguments");
  if (n == 0) return PF_OK;
  hipStream_t st = as_stream(stream);
  if (reinterpret_cast<uintptr_t>(data) % 16 == 0 && n >= 4) {
    const int64_t n4 = n / 4;
    const unsigned grid = (unsigned)std::min<int64_t>((n4 + 255) / 256, 4096);
    scrub4_kernel<<<grid, 256, 0, st>>>(reinterpret_cast<float4*>(data), n4);
    if (n % 4) scrub_kernel<<<1, 256, 0, st>>>(data + 4 * n4, n % 4);
  } else {
    scrub_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(data, n);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
