// Tile binning of power-cell screen footprints.
//
// Replaces, per frame:
//   count_visible_pinhole_kernel  rasterize.py:140-200
//   torch.cumsum + .item()        rasterize.py:1713-1714
//   write_visible_pinhole_kernel  rasterize.py:293-363
//   torch.argsort + gather        rasterize.py:1761-1768
// and for ray-map cameras
//   _precompute_cones, compute_leaf_cones_kernel, merge_cones_kernel
//                                 rasterize.py:1518-1585, 468-569
//   count/write_visible_generic_kernel  rasterize.py:202-291, 365-466
//
// This file is compiled with -ffp-contract=off: tile membership and sort keys
// are float-derived integers and must not depend on FMA contraction. Every
// float operation below is a single correctly rounded IEEE operation, in the
// order the reference writes it (camera.py:79-178, rasterize.py:155-190).
//
// Sorting: the count pass yields per-tile offsets, so the reference's global
// 64-bit argsort becomes a sort inside each tile segment. Each segment holds
// 64-bit records (depth key << 32 | site); one 256-thread workgroup sorts one
// segment with an ascending-only bitonic network (virtual +inf padding to the
// next power of two), staged in LDS up to 4096 records and merged through
// global memory above that. Equal depth keys are ordered by site id, so the
// lists are deterministic (the reference's argsort is not stable).
#include <stdarg.h>
#include <stdio.h>

#include <mutex>

#include "pf_common.hpp"

namespace pf {

struct BinCam {
  f3 eye, right, up;
  int W, H, th;
};

static BinCam make_bincam(const pf_camera& c) {
  BinCam b;
  b.eye = f3{c.eye[0], c.eye[1], c.eye[2]};
  b.right = f3{c.right[0], c.right[1], c.right[2]};
  b.up = f3{c.up[0], c.up[1], c.up[2]};
  b.W = c.width;
  b.H = c.height;
  b.th = tiles_h(c);
  return b;
}

// float -> int32 as the hardware converts: truncate, saturate, NaN -> 0
__device__ __forceinline__ int f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int)0x80000000;
  return (int)x;
}

struct Obb {
  float uc, vc, r1, r2, vx, vy;
};

// camera.py:79-128
__device__ __forceinline__ Obb proj_sphere_to_obb(const BinCam& cam, f3 center, float radius) {
  Obb o;
  f3 X = cam.right, Y = cam.up;
  f3 C = center - cam.eye;
  float C2 = dot(C, C);
  float R2 = radius * radius;
  f3 zc = cross(Y, X);
  float zl = sqrtf(dot(zc, zc));
  f3 Z = zl > 0.0f ? f3{zc.x / zl, zc.y / zl, zc.z / zl} : f3{0.0f, 0.0f, 0.0f};
  float cx = dot(C, X), cy = dot(C, Y), cz = dot(C, Z);
  float k = C2 - R2;
  float a = cx * cx - k * dot(X, X);
  float b = cy * cy - k * dot(Y, Y);
  float f = cz * cz - k * dot(Z, Z);
  float c = 2.0f * cx * cy;
  float d = 2.0f * cx * cz;
  float e = 2.0f * cy * cz;
  float denom = 4.0f * a * b - c * c;
  if (fabsf(denom) < 1e-8f) {
    o.uc = 0.0f; o.vc = 0.0f; o.r1 = 0.0f; o.r2 = 0.0f; o.vx = 1.0f; o.vy = 0.0f;
    return o;
  }
  float uc = (c * e - 2.0f * b * d) / denom;
  float vc = (c * d - 2.0f * a * e) / denom;
  float F0 = a * uc * uc + b * vc * vc + c * uc * vc + d * uc + e * vc + f;
  float ts = sqrtf((a - b) * (a - b) + c * c);
  float l1 = (a + b + ts) / 2.0f;
  float l2 = (a + b - ts) / 2.0f;
  float r1 = sqrtf(-F0 / l1);
  float r2 = sqrtf(-F0 / l2);
  float vx = 0.5f * c;
  float vy = l1 - a;
  float norm = sqrtf(vx * vx + vy * vy);
  if (norm < 1e-6f) {
    vx = 1.0f; vy = 0.0f;
  } else {
    vx = vx / norm; vy = vy / norm;
  }
  o.uc = uc; o.vc = vc; o.r1 = r1; o.r2 = r2; o.vx = vx; o.vy = vy;
  return o;
}

// camera.py:131-178 (two OBB axes)
__device__ __forceinline__ bool tile_obb_test(int tile_i, int tile_j, const BinCam& cam, const Obb& o) {
  float wm1 = (float)(cam.W - 1), hm1 = (float)(cam.H - 1);
  float u_min = 2.0f * (float)(tile_j * PF_TILE_WIDTH) / wm1 - 1.0f;
  float u_max = 2.0f * (float)((tile_j + 1) * PF_TILE_WIDTH) / wm1 - 1.0f;
  float v_max = 1.0f - 2.0f * (float)(tile_i * PF_TILE_WIDTH) / hm1;
  float v_min = 1.0f - 2.0f * (float)((tile_i + 1) * PF_TILE_WIDTH) / hm1;
  float u_center = (u_min + u_max) * 0.5f;
  float v_center = (v_min + v_max) * 0.5f;
  float eu = (u_max - u_min) * 0.5f;
  float ev = (v_max - v_min) * 0.5f;
  float du = u_center - o.uc;
  float dv = v_center - o.vc;
  float p1 = du * o.vx + dv * o.vy;
  float t1 = eu * fabsf(o.vx) + ev * fabsf(o.vy);
  if (fabsf(p1) > o.r1 + t1) return false;
  float p2 = du * -o.vy + dv * o.vx;
  float t2 = eu * fabsf(-o.vy) + ev * fabsf(o.vx);
  if (fabsf(p2) > o.r2 + t2) return false;
  return true;
}

struct Footprint {
  int ti0, ti1, tj0, tj1;
  Obb obb;
  uint32_t key;
};

// rasterize.py:150-190, 360-363
__device__ __forceinline__ bool site_footprint(const BinCam& cam, float4 sph, Footprint& fp) {
  f3 center{sph.x, sph.y, sph.z};
  float radius = sph.w;
  f3 v = center - cam.eye;
  f3 fwd = cross(cam.up, cam.right);  // not normalised, as the reference
  if (sqrtf(dot(v, v)) < 4.0f * radius || dot(v, fwd) < 0.1f) return false;
  Obb o = proj_sphere_to_obb(cam, center, radius);
  float ext_u = fabsf(o.vx * o.r1) + fabsf(o.vy * o.r2);
  float ext_v = fabsf(o.vy * o.r1) + fabsf(o.vx * o.r2);
  float u_min = o.uc - ext_u, u_max = o.uc + ext_u;
  float v_min = o.vc - ext_v, v_max = o.vc + ext_v;
  float wm1 = (float)(cam.W - 1), hm1 = (float)(cam.H - 1);
  float j_min_f = (u_min + 1.0f) * 0.5f * wm1;
  float j_max_f = (u_max + 1.0f) * 0.5f * wm1;
  float i_min_f = (1.0f - v_max) * 0.5f * hm1;
  float i_max_f = (1.0f - v_min) * 0.5f * hm1;
  int i_min = max(f2i(floorf(i_min_f)), 0);
  int i_max = min(f2i(ceilf(i_max_f)), cam.H - 1);
  int j_min = max(f2i(floorf(j_min_f)), 0);
  int j_max = min(f2i(ceilf(j_max_f)), cam.W - 1);
  fp.ti0 = i_min / PF_TILE_WIDTH;
  fp.ti1 = i_max / PF_TILE_WIDTH;
  fp.tj0 = j_min / PF_TILE_WIDTH;
  fp.tj1 = j_max / PF_TILE_WIDTH;
  fp.obb = o;
  float pow_dist = dot(v, v) - radius * radius;
  fp.key = __builtin_bit_cast(uint32_t, pow_dist);
  return true;
}

__global__ void __launch_bounds__(256) fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Candidate tiles of a wave's sites, dealt out to its lanes. One lane per site
// leaves a wave as slow as its largest footprint (a site near the camera spans
// thousands of tiles, the median a handful), so each lane computes its site's
// tile rectangle, the wave scans the rectangle sizes, and the lanes then take
// the wave's candidates 64 at a time; a lane finds the site owning candidate k
// by binary search over the scanned starts in LDS. Which lane tests which
// (site, tile) pair changes only the order of the atomics below, and the
// per-tile sort removes that order.
struct WaveSite {
  float uc, vc, r1, r2, vx, vy;
  Footprint ball;  // footprint of the site's bound (pf_live_sites), when has_ball
  int ti0, tj0, nj, start;
  uint32_t key;
  int site, has_ball;
};

// A tile in a footprint: inside its tile rectangle and through the OBB test,
// as binning decides for a sphere.
__device__ __forceinline__ bool in_footprint(int ti, int tj, const BinCam& cam, const Footprint& f) {
  return ti >= f.ti0 && ti <= f.ti1 && tj >= f.tj0 && tj <= f.tj1 && tile_obb_test(ti, tj, cam, f.obb);
}

// f(tile, site record, in_ball): in_ball is false when the tile misses the
// site's bound ball (no ray of the tile can be shaded by the site).
template <class F>
__device__ __forceinline__ void for_each_overlap(const BinCam& cam, WaveSite* ws, int lane, bool valid,
                                                 const Footprint& fp, int site, const uint8_t* live,
                                                 const float4* bounds, F&& f) {
  const int ni = valid ? max(fp.ti1 - fp.ti0 + 1, 0) : 0;
  const int nj = valid ? max(fp.tj1 - fp.tj0 + 1, 0) : 0;
  const int nc = ni * nj;
  int incl = nc;
#pragma unroll
  for (int m = 1; m < WAVE; m <<= 1) {
    const int v = __shfl_up(incl, m, WAVE);
    if (lane >= m) incl += v;
  }
  const int total = __shfl(incl, WAVE - 1, WAVE);
  wave_sync();
  WaveSite& me = ws[lane];
  me.uc = fp.obb.uc;
  me.vc = fp.obb.vc;
  me.r1 = fp.obb.r1;
  me.r2 = fp.obb.r2;
  me.vx = fp.obb.vx;
  me.vy = fp.obb.vy;
  me.ti0 = fp.ti0;
  me.tj0 = fp.tj0;
  me.nj = nj;
  me.start = incl - nc;
  me.key = fp.key;
  me.site = site;
  // a ball that fails the sphere cull does not restrict its site
  me.has_ball = (valid && bounds && (!live || live[site]) && site_footprint(cam, bounds[site], me.ball)) ? 1 : 0;
  wave_sync();
  for (int base = 0; base < total; base += WAVE) {
    const int k = base + lane;
    if (k < total) {
      // largest l with start[l] <= k; it owns k because start[l + 1] > k
      int l = 0;
#pragma unroll
      for (int step = WAVE / 2; step >= 1; step >>= 1)
        if (ws[l + step].start <= k) l += step;
      const WaveSite& o = ws[l];
      const int local = k - o.start;
      const int ti = o.ti0 + local / o.nj, tj = o.tj0 + local % o.nj;
      const Obb obb{o.uc, o.vc, o.r1, o.r2, o.vx, o.vy};
      if (tile_obb_test(ti, tj, cam, obb)) f(ti + tj * cam.th, o, !o.has_ball || in_footprint(ti, tj, cam, o.ball));
    }
  }
}

__global__ void __launch_bounds__(256)
    bin_count_kernel(BinCam cam, const float4* __restrict__ spheres, const uint8_t* __restrict__ live,
                     const float4* __restrict__ bounds, int n, int32_t* __restrict__ counts) {
  __shared__ WaveSite s_ws[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (WAVE - 1);
  Footprint fp{};
  const bool valid = s < n && (!live || live[s]) && site_footprint(cam, spheres[s], fp);
  for_each_overlap(cam, s_ws + (threadIdx.x & ~(WAVE - 1)), lane, valid, fp, s, live, bounds,
                   [&](int t, const WaveSite&, bool in_ball) {
                     if (in_ball) atomicAdd(&counts[t + 1], 1);
                   });
}

__global__ void __launch_bounds__(256)
    bin_emit_kernel(BinCam cam, const float4* __restrict__ spheres, const uint8_t* __restrict__ live,
                    const float4* __restrict__ bounds, int n, const int64_t* __restrict__ offsets,
                    uint64_t* __restrict__ records, int32_t* __restrict__ cursor, uint8_t* __restrict__ visible,
                    int32_t* __restrict__ dropped) {
  __shared__ WaveSite s_ws[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (WAVE - 1);
  Footprint fp{};
  // a site left out of the lists still walks its tiles when the visibility
  // mask is asked for: the mask is the footprint's, as without the mask
  const bool valid = s < n && (visible || !live || live[s]) && site_footprint(cam, spheres[s], fp);
  for_each_overlap(cam, s_ws + (threadIdx.x & ~(WAVE - 1)), lane, valid, fp, s, live, bounds,
                   [&](int t, const WaveSite& o, bool in_ball) {
    if (in_ball && (!live || live[o.site])) {
      const int64_t idx = offsets[t] + (int64_t)atomicAdd(&cursor[t], 1);
      if (idx < offsets[t + 1])
        records[idx] = ((uint64_t)o.key << 32) | (uint32_t)o.site;
      else if (dropped)
        atomicAdd(dropped, 1);  // count and emit disagree: reported, not silent
    }
    if (visible) visible[o.site] = 1;
  });
}

// Sites that cannot shade a ray from the camera. A site's shaded interval is
// its ray-sphere interval clipped by the planes of its power faces; when the
// whole sphere lies beyond one of those planes, every ray's clipped interval
// is empty and the compositing kernels give the site no contribution and no
// gradient, so the tile lists can leave it out. The plane is the one those
// kernels evaluate: the decoded fp16 deltas (radius form of pf_adj_encode,
// power-weight form of the benchmark) or full-precision spheres. The sphere
// has to clear the plane by more than the rounding of the kernels' interval
// ends: a tangent ray's ends move by about 5e-4 sqrt(r |c - eye|) (the square
// root of a rounded r^2 - |oc x d|^2), a plane crossing by a few ulp of the
// coordinates involved; the margin is 16 and about 170 times these.
// A wave takes 64 consecutive sites and deals their edges to its lanes, as
// the adjacency encoder does.
__global__ void __launch_bounds__(256)
    live_sites_kernel(const float4* __restrict__ spheres, int n, const int32_t* __restrict__ adj,
                      const int32_t* __restrict__ off, int64_t n_edges, const uint2* __restrict__ diff, int power,
                      f3 eye, uint8_t* __restrict__ live) {
  __shared__ int s_off[4][WAVE + 1];
  __shared__ float4 s_sph[4][WAVE];
  __shared__ int s_dead[4][WAVE];
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int s0 = (blockIdx.x * 4 + wv) * WAVE;
  if (s0 >= n) return;
  const int ns = min(WAVE, n - s0);
  int* so = s_off[wv];
  float4* sp = s_sph[wv];
  int* sd = s_dead[wv];
  if (lane < ns) {
    sp[lane] = spheres[s0 + lane];
    so[lane] = off[s0 + lane];
  }
  if (lane == 0) so[ns] = off[s0 + ns];  // ns may be 64: one past the last lane
  sd[lane] = 0;
  wave_sync();
  const float eye_len = __builtin_sqrtf(eye.x * eye.x + eye.y * eye.y + eye.z * eye.z);
  const int e1 = (int)min((int64_t)so[ns], n_edges);  // an inconsistent CSR ends at the edge arrays
  for (int e = so[0] + lane; e < e1; e += WAVE) {
    int l = 0;
#pragma unroll
    for (int step = WAVE / 2; step >= 1; step >>= 1)
      if (l + step < ns && so[l + step] <= e) l += step;
    const float4 c = sp[l];
    f3 nv;
    float h;
    if (diff) {
      const uint2 raw = diff[e];
      nv = f3{half_bits_to_float((uint16_t)(raw.x & 0xffff)), half_bits_to_float((uint16_t)(raw.x >> 16)),
              half_bits_to_float((uint16_t)(raw.y & 0xffff))};
      const float q = half_bits_to_float((uint16_t)(raw.y >> 16));
      h = power ? q - (c.x * nv.x + c.y * nv.y + c.z * nv.z) : 0.5f * (dot(nv, nv) + c.w * c.w - q * q);
    } else {
      const float4 a = spheres[adj[e]];
      nv = f3{a.x - c.x, a.y - c.y, a.z - c.z};
      h = 0.5f * (dot(nv, nv) + c.w * c.w - a.w * a.w);
    }
    // the cell is {y : y.n <= h} around the centre; the sphere reaches
    // y.n >= -r |n| only, so it lies beyond by G = -h / |n| - r
    const float nl = __builtin_sqrtf(dot(nv, nv));
    const f3 ce{c.x - eye.x, c.y - eye.y, c.z - eye.z};
    const float dist = __builtin_sqrtf(dot(ce, ce)) + c.w;
    const float c_len = __builtin_sqrtf(c.x * c.x + c.y * c.y + c.z * c.z);
    const float m = 8e-3f * __builtin_sqrtf(c.w * dist) + 1e-5f * (dist + c_len + eye_len);
    if (nl > 0.0f && -h - c.w * nl > m * nl + 1e-5f * fabsf(h)) sd[l] = 1;
  }
  wave_sync();
  if (lane < ns) live[s0 + lane] = sd[lane] ? 0 : 1;
}

// A ball enclosing every point a site can shade. It starts as the sphere,
// inflated by 2e-3 sqrt(r |c - eye|) (four times the movement of a tangent
// ray's interval ends under FP32 rounding); each face whose plane, moved
// outward by the plane-rounding margin of live_sites_kernel, leaves the
// ball's centre outside replaces the ball by the smallest ball enclosing the
// cap on the cell's side (centre on the plane, radius sqrt(rho^2 - delta^2)).
// Every point the site shades lies in each cap, so it stays inside; three
// sweeps over the faces in list order. A cap that vanishes (the ball wholly
// outside a plane) leaves nothing to shade: the site is marked not live.
// Binning then keeps a (tile, site) pair only if the tile also meets the
// ball's footprint. The sweeps are sequential, so a lane takes a site; a
// wave first gathers the live sites among its 256 (most sites are not) so
// that its lanes stay busy. A site that is not live keeps its sphere.
constexpr int BOUNDS_SPAN = 4 * WAVE;  // sites per wave

__device__ __forceinline__ void site_bound(const float4* __restrict__ spheres, int s,
                                           const int32_t* __restrict__ adj, const int32_t* __restrict__ off,
                                           int64_t n_edges, const uint2* __restrict__ diff, int power, f3 eye,
                                           uint8_t* __restrict__ live, float4* __restrict__ bounds) {
  const float4 c = spheres[s];
  const f3 ce{c.x - eye.x, c.y - eye.y, c.z - eye.z};
  const float dist = __builtin_sqrtf(dot(ce, ce)) + c.w;
  const float c_len = __builtin_sqrtf(c.x * c.x + c.y * c.y + c.z * c.z);
  const float eye_len = __builtin_sqrtf(eye.x * eye.x + eye.y * eye.y + eye.z * eye.z);
  f3 b{c.x, c.y, c.z};
  float rad = c.w + 2e-3f * __builtin_sqrtf(c.w * dist);
  bool empty = false;
  const int e0 = (int)min((int64_t)off[s], n_edges), e1 = (int)min((int64_t)off[s + 1], n_edges);
  // faces are read four at a time (the sweeps are a chain of dependent
  // steps; one load per step would leave every step waiting on memory)
  for (int pass = 0; pass < 3 && !empty; pass++) {
    for (int e = e0; e < e1 && !empty; e += 4) {
      float4 raw[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        raw[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (e + j < e1) {
          if (diff) {
            const uint2 r2 = diff[e + j];
            raw[j] = make_float4(__builtin_bit_cast(float, r2.x), __builtin_bit_cast(float, r2.y), 0.0f, 0.0f);
          } else {
            raw[j] = spheres[adj[e + j]];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (e + j >= e1 || empty) break;
        f3 nv;
        float h;
        if (diff) {
          const uint32_t bx = __builtin_bit_cast(uint32_t, raw[j].x), by = __builtin_bit_cast(uint32_t, raw[j].y);
          nv = f3{half_bits_to_float((uint16_t)(bx & 0xffff)), half_bits_to_float((uint16_t)(bx >> 16)),
                  half_bits_to_float((uint16_t)(by & 0xffff))};
          const float q = half_bits_to_float((uint16_t)(by >> 16));
          h = power ? q - (c.x * nv.x + c.y * nv.y + c.z * nv.z) : 0.5f * (dot(nv, nv) + c.w * c.w - q * q);
        } else {
          const float4 a = raw[j];
          nv = f3{a.x - c.x, a.y - c.y, a.z - c.z};
          h = 0.5f * (dot(nv, nv) + c.w * c.w - a.w * a.w);
        }
        const float n2 = dot(nv, nv);
        if (!(n2 > 0.0f)) continue;
        // hardware 1/sqrt (about 1 ulp): the ball's margins are orders of
        // magnitude wider than that rounding
        const float inv = __builtin_amdgcn_rsqf(n2);
        const f3 u{nv.x * inv, nv.y * inv, nv.z * inv};
        const float m = 1e-5f * (dist + c_len + eye_len + fabsf(h) * inv);
        const float hd = h * inv + m;
        const f3 bc{b.x - c.x, b.y - c.y, b.z - c.z};
        const float delta = hd - dot(bc, u);
        if (delta <= -rad) {
          empty = true;
        } else if (delta < 0.0f) {
          b = f3{b.x + u.x * delta, b.y + u.y * delta, b.z + u.z * delta};
          rad = __builtin_sqrtf(fmaxf(rad * rad - delta * delta, 0.0f));
        }
      }
    }
  }
  if (empty) live[s] = 0;
  bounds[s] = make_float4(b.x, b.y, b.z, rad);
}

__global__ void __launch_bounds__(256)
    site_bounds_kernel(const float4* __restrict__ spheres, int n, const int32_t* __restrict__ adj,
                       const int32_t* __restrict__ off, int64_t n_edges, const uint2* __restrict__ diff, int power,
                       f3 eye, uint8_t* __restrict__ live, float4* __restrict__ bounds) {
  __shared__ int s_idx[4][BOUNDS_SPAN];
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int s0 = (blockIdx.x * 4 + wv) * BOUNDS_SPAN;
  if (s0 >= n) return;
  int* idx = s_idx[wv];
  int cnt = 0;
  for (int g = 0; g < BOUNDS_SPAN; g += WAVE) {
    const int s = s0 + g + lane;
    bool is_live = false;
    if (s < n) {
      is_live = live[s] != 0;
      if (!is_live) bounds[s] = spheres[s];
    }
    const uint64_t m = __ballot(is_live);
    if (is_live) idx[cnt + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = s;
    cnt += __popcll(m);
  }
  wave_sync();
  for (int i = lane; i < cnt; i += WAVE) site_bound(spheres, idx[i], adj, off, n_edges, diff, power, eye, live, bounds);
}

// ------------------------------------------------ non-pinhole cameras
// Ray-map cameras bin through a quadtree of ray cones (rasterize.py:1518-1585):
// level 0 holds one cone per 2x2 pixel block, each further level one per 2x2
// block of the level below (dims rounded up) until 1x1, level after level,
// row-major inside a level; level 2 is exactly the 8x8 tiles. A cone is
// (unit axis, cos of the half-angle); cos > 1 marks a node without rays.
//
// A site is binned to every tile whose cone, and every ancestor's cone, meets
// its ball (sphere_cone_intersect, camera.py:27-44). The reference descends
// with a 32-entry stack per thread; here the descent is stackless: a quadtree
// over a grid has implicit parent/child/sibling links, so the walk needs only
// (level, row, column) in registers. Visiting order differs from the
// reference's; it only decides atomic slot order, which the per-tile sort
// removes.
constexpr int CONE_LEAF = 2;
constexpr int CONE_MAX_LEVELS = 14;  // images up to 8192 pixels a side

struct ConeTree {
  int n_levels, tile_level, total;
  int off[CONE_MAX_LEVELS], dh[CONE_MAX_LEVELS], dw[CONE_MAX_LEVELS];
};

static bool make_cone_tree(int W, int H, ConeTree& t) {
  t = ConeTree{};
  int dh = (H + CONE_LEAF - 1) / CONE_LEAF, dw = (W + CONE_LEAF - 1) / CONE_LEAF;
  t.n_levels = 0;
  t.total = 0;
  for (;;) {
    if (t.n_levels == CONE_MAX_LEVELS) return false;
    t.off[t.n_levels] = t.total;
    t.dh[t.n_levels] = dh;
    t.dw[t.n_levels] = dw;
    t.total += dh * dw;
    t.n_levels++;
    if (dh == 1 && dw == 1) break;
    dh = (dh + 1) / 2;
    dw = (dw + 1) / 2;
  }
  t.tile_level = 0;
  for (int sz = CONE_LEAF; sz < PF_TILE_WIDTH; sz *= 2) t.tile_level++;
  return t.n_levels > t.tile_level;
}

// compute_leaf_cones_kernel (rasterize.py:468-512): one thread per leaf.
__global__ void __launch_bounds__(256)
    cone_leaf_kernel(const float* __restrict__ ray_maps, int W, int H, int dh, int dw, float4* __restrict__ cones) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= dh * dw) return;
  const int a = id / dw, b = id % dw;
  f3 rays[4];
  f3 sum{0.0f, 0.0f, 0.0f};
  int nv = 0;
#pragma unroll
  for (int di = 0; di < CONE_LEAF; di++)
#pragma unroll
    for (int dj = 0; dj < CONE_LEAF; dj++) {
      const int pi = min(a * CONE_LEAF + di, H - 1), pj = min(b * CONE_LEAF + dj, W - 1);
      const float* rm = ray_maps + ((size_t)pi * W + pj) * 6;
      const f3 r{rm[3], rm[4], rm[5]};
      const float len = sqrtf(dot(r, r));
      if (len > 1.0e-6f) {
        const f3 rn{r.x / len, r.y / len, r.z / len};
        rays[nv++] = rn;
        sum = sum + rn;
      }
    }
  if (nv == 0) {
    cones[id] = make_float4(0.0f, 0.0f, 0.0f, 2.0f);
    return;
  }
  const float sl = sqrtf(dot(sum, sum));
  const f3 axis = sl > 0.0f ? f3{sum.x / sl, sum.y / sl, sum.z / sl} : f3{0.0f, 0.0f, 0.0f};
  float ch = 1.0f;
  for (int k = 0; k < nv; k++) {
    const float d = dot(axis, rays[k]);
    ch = d < ch ? d : ch;
  }
  cones[id] = make_float4(axis.x, axis.y, axis.z, ch);
}

// merge_cones_kernel (rasterize.py:514-569): one thread per parent node.
__global__ void __launch_bounds__(256) cone_merge_kernel(float4* __restrict__ cones, int poff, int pdh, int pdw,
                                                         int coff, int cdh, int cdw) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= pdh * pdw) return;
  const int a = id / pdw, b = id % pdw;
  // children in (0,0) (0,1) (1,0) (1,1) order; fixed slots keep them in registers
  float4 c[4];
  bool ok[4];
  f3 sum{0.0f, 0.0f, 0.0f};
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int ca = 2 * a + (k >> 1), cb = 2 * b + (k & 1);
    ok[k] = false;
    c[k] = make_float4(0.0f, 0.0f, 0.0f, 2.0f);
    if (ca < cdh && cb < cdw) {
      c[k] = cones[coff + ca * cdw + cb];
      ok[k] = c[k].w <= 1.0f;
    }
    if (ok[k]) {
      sum = sum + f3{c[k].x, c[k].y, c[k].z};
      any = true;
    }
  }
  if (!any) {
    cones[poff + id] = make_float4(0.0f, 0.0f, 0.0f, 2.0f);
    return;
  }
  const float sl = sqrtf(dot(sum, sum));
  const f3 axis = sl > 0.0f ? f3{sum.x / sl, sum.y / sl, sum.z / sl} : f3{0.0f, 0.0f, 0.0f};
  float ch = 1.0f;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (!ok[k]) continue;
    const float d = dot(axis, f3{c[k].x, c[k].y, c[k].z});
    const float sa = 1.0f - d * d, sc = 1.0f - c[k].w * c[k].w;
    const float ext = d * c[k].w - sqrtf(sa > 0.0f ? sa : 0.0f) * sqrtf(sc > 0.0f ? sc : 0.0f);
    ch = ext < ch ? ext : ch;
  }
  cones[poff + id] = make_float4(axis.x, axis.y, axis.z, ch);
}

// Per-site part of sphere_cone_intersect (camera.py:27-44) and the cull of
// rasterize.py:226-229.
struct ConeSite {
  f3 v;
  float dist, sin_beta, cos_beta;
  uint32_t key;
};

__device__ __forceinline__ bool cone_site(const BinCam& cam, float4 sph, ConeSite& cs) {
  const f3 center{sph.x, sph.y, sph.z};
  const float radius = sph.w;
  cs.v = center - cam.eye;
  const f3 fwd = cross(cam.up, cam.right);
  cs.dist = sqrtf(dot(cs.v, cs.v));
  if (cs.dist < 4.0f * radius || dot(cs.v, fwd) < 0.1f) return false;
  cs.sin_beta = radius / cs.dist;
  cs.cos_beta = sqrtf(1.0f - cs.sin_beta * cs.sin_beta);
  const float pow_dist = dot(cs.v, cs.v) - radius * radius;
  cs.key = __builtin_bit_cast(uint32_t, pow_dist);
  return true;
}

__device__ __forceinline__ bool cone_hit(const ConeSite& cs, float4 cone) {
  const float cos_alpha = dot(cs.v, f3{cone.x, cone.y, cone.z}) / cs.dist;
  const float sin_half = sqrtf(1.0f - cone.w * cone.w);
  const float threshold = cone.w * cs.cos_beta - sin_half * cs.sin_beta;
  return cos_alpha > threshold;
}

// Stackless depth-first walk; visit(tile) for every tile reached. The level
// tables come from LDS (the level index differs per lane).
template <typename Visit>
__device__ __forceinline__ void cone_walk(const ConeSite& cs, const float4* __restrict__ cones, const int* s_off,
                                          const int* s_dh, const int* s_dw, int top, int tile_level, int th,
                                          Visit visit) {
  int level = top, a = 0, b = 0;
  for (;;) {
    const float4 cone = cones[s_off[level] + a * s_dw[level] + b];
    if (cone.w <= 1.0f && cone_hit(cs, cone)) {
      if (level == tile_level) {
        visit(a + b * th);
      } else {
        level--;
        a *= 2;
        b *= 2;
        continue;  // the (0,0) child always exists
      }
    }
    // next sibling in (0,0) (0,1) (1,0) (1,1) order, else climb
    for (;;) {
      if (level == top) return;
      if (!(b & 1) && b + 1 < s_dw[level]) {
        b++;
        break;
      }
      if (!(a & 1) && a + 1 < s_dh[level]) {
        a++;
        b &= ~1;
        break;
      }
      level++;
      a >>= 1;
      b >>= 1;
    }
  }
}

#define CONE_LDS_TABLES(t)                                                            \
  __shared__ int s_off[CONE_MAX_LEVELS], s_dh[CONE_MAX_LEVELS], s_dw[CONE_MAX_LEVELS]; \
  if (threadIdx.x < CONE_MAX_LEVELS) {                                                \
    s_off[threadIdx.x] = t.off[threadIdx.x];                                          \
    s_dh[threadIdx.x] = t.dh[threadIdx.x];                                            \
    s_dw[threadIdx.x] = t.dw[threadIdx.x];                                            \
  }                                                                                   \
  __syncthreads()

// count_visible_generic_kernel (rasterize.py:202-291)
__global__ void __launch_bounds__(256)
    bin_count_cones_kernel(BinCam cam, ConeTree t, const float4* __restrict__ cones,
                           const float4* __restrict__ spheres, int n, int32_t* __restrict__ counts) {
  CONE_LDS_TABLES(t);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  ConeSite cs;
  if (!cone_site(cam, spheres[s], cs)) return;
  cone_walk(cs, cones, s_off, s_dh, s_dw, t.n_levels - 1, t.tile_level, cam.th,
            [&](int tile) { atomicAdd(&counts[tile + 1], 1); });
}

// write_visible_generic_kernel (rasterize.py:365-466)
__global__ void __launch_bounds__(256)
    bin_emit_cones_kernel(BinCam cam, ConeTree t, const float4* __restrict__ cones,
                          const float4* __restrict__ spheres, int n, const int64_t* __restrict__ offsets,
                          uint64_t* __restrict__ records, int32_t* __restrict__ cursor,
                          uint8_t* __restrict__ visible, int32_t* __restrict__ dropped) {
  CONE_LDS_TABLES(t);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  ConeSite cs;
  if (!cone_site(cam, spheres[s], cs)) return;
  const uint64_t rec = ((uint64_t)cs.key << 32) | (uint32_t)s;
  bool any = false;
  cone_walk(cs, cones, s_off, s_dh, s_dw, t.n_levels - 1, t.tile_level, cam.th, [&](int tile) {
    const int64_t idx = offsets[tile] + (int64_t)atomicAdd(&cursor[tile], 1);
    if (idx < offsets[tile + 1]) {
      records[idx] = rec;
      any = true;
    } else if (dropped) {
      atomicAdd(dropped, 1);
    }
  });
  if (any && visible) visible[s] = 1;
}

// Inclusive scan of counts[0..n] (int32) into offsets[0..n] (int64) by one
// 1024-thread workgroup: each thread scans a contiguous run serially, then
// the run totals are scanned across the workgroup.
__global__ void __launch_bounds__(1024)
    bin_scan_kernel(const int32_t* __restrict__ counts, int64_t n, int64_t* __restrict__ offsets) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t a = t * per;
  const int64_t b = min(a + per, n);
  int64_t s = 0;
  for (int64_t i = a; i < b; i++) s += counts[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = t > 0 ? part[t - 1] : 0;
  for (int64_t i = a; i < b; i++) {
    run += counts[i];
    offsets[i] = run;
  }
}

constexpr int SORT_THREADS = 256;
constexpr int SORT_CAP = 4096;  // records staged in LDS (32 KiB)

// One step of the ascending-only bitonic network over n valid records in a
// power-of-two frame: either the mirrored first step of a merge of size k
// (mirror = true) or a half-cleaner with stride j. Pairs whose upper index
// is >= n compare against virtual +inf and are skipped. k and j are powers
// of two; lg_half = log2(k/2) for the mirror step, lg_j = log2(j) otherwise.
template <int THREADS, typename Buf>
__device__ __forceinline__ void bitonic_step(Buf* buf, int64_t n_pairs, int64_t n, int lg, bool mirror,
                                             int tid = -1) {
  const int64_t m = ((int64_t)1 << lg) - 1;
  for (int64_t p = tid >= 0 ? tid : (int)threadIdx.x; p < n_pairs; p += THREADS) {
    const int64_t blk = p >> lg, l = p & m;
    int64_t i, q;
    if (mirror) {  // block of size k = 2^(lg+1): pair l with k-1-l
      i = (blk << (lg + 1)) + l;
      q = (blk << (lg + 1)) + (m << 1) + 1 - l;
    } else {  // stride j = 2^lg
      i = (blk << (lg + 1)) + l;
      q = i + m + 1;
    }
    if (q >= n) continue;
    uint64_t a = buf[i], b = buf[q];
    if (a > b) {
      buf[i] = b;
      buf[q] = a;
    }
  }
}

__device__ __forceinline__ int ilog2(int64_t x) { return 63 - __builtin_clzll((unsigned long long)x); }

// Sorts tile `tile`'s segment with CAP records staged in LDS by THREADS
// threads; longer segments merge through global memory.
template <int CAP, int THREADS>
__device__ void sort_tile(const int64_t* __restrict__ offsets, int tile, uint64_t* __restrict__ records,
                          int32_t* __restrict__ sites, uint64_t* lds) {
  const int64_t a = offsets[tile];
  const int64_t n = offsets[tile + 1] - a;
  if (n <= 0) return;
  uint64_t* g = records + a;
  if (n == 1) {
    if (threadIdx.x == 0) sites[a] = (int32_t)(uint32_t)g[0];
    return;
  }
  int64_t P = 1;
  while (P < n) P <<= 1;
  const int lgP = ilog2(P);
  if (n <= CAP) {
    for (int64_t i = threadIdx.x; i < n; i += THREADS) lds[i] = g[i];
    __syncthreads();
    for (int lk = 1; lk <= lgP; lk++) {
      bitonic_step<THREADS>(lds, P / 2, n, lk - 1, true);
      __syncthreads();
      for (int lj = lk - 2; lj >= 0; lj--) {
        bitonic_step<THREADS>(lds, P / 2, n, lj, false);
        __syncthreads();
      }
    }
    for (int64_t i = threadIdx.x; i < n; i += THREADS) sites[a + i] = (int32_t)(uint32_t)lds[i];
    return;
  }
  // Large segment: sort CAP chunks in LDS, then merge through global
  // memory; strides that fit a chunk are finished in LDS.
  const int lgC = ilog2(CAP);
  const int64_t n_chunks = (n + CAP - 1) / CAP;
  for (int lk = 1; lk <= lgP; lk++) {
    if (lk > lgC) {
      bitonic_step<THREADS>(g, P / 2, n, lk - 1, true);
      __syncthreads();
      for (int lj = lk - 2; lj >= lgC; lj--) {
        bitonic_step<THREADS>(g, P / 2, n, lj, false);
        __syncthreads();
      }
    }
    if (lk > lgC || lk == lgC) {
      // finish every step of merge sizes <= 2^lk whose pairs stay in a chunk
      for (int64_t c = 0; c < n_chunks; c++) {
        const int64_t base = c * CAP;
        const int64_t cn = min((int64_t)CAP, n - base);
        for (int64_t i = threadIdx.x; i < cn; i += THREADS) lds[i] = g[base + i];
        __syncthreads();
        if (lk == lgC) {
          for (int l2 = 1; l2 <= lgC; l2++) {
            bitonic_step<THREADS>(lds, CAP / 2, cn, l2 - 1, true);
            __syncthreads();
            for (int lj = l2 - 2; lj >= 0; lj--) {
              bitonic_step<THREADS>(lds, CAP / 2, cn, lj, false);
              __syncthreads();
            }
          }
        } else {
          for (int lj = lgC - 1; lj >= 0; lj--) {
            bitonic_step<THREADS>(lds, CAP / 2, cn, lj, false);
            __syncthreads();
          }
        }
        for (int64_t i = threadIdx.x; i < cn; i += THREADS) g[base + i] = lds[i];
        __syncthreads();
      }
    }
  }
  for (int64_t i = threadIdx.x; i < n; i += THREADS) sites[a + i] = (int32_t)(uint32_t)g[i];
}

// One workgroup per tile. With `skip_big`, tiles above SORT_CAP are left to
// bin_sort_big_kernel; with `skip_small`, tiles of at most SMALL_CAP records
// to bin_sort_small_kernel.
constexpr int SMALL_CAP = 512;

__global__ void __launch_bounds__(SORT_THREADS)
    bin_sort_kernel(const int64_t* __restrict__ offsets, int n_tiles, uint64_t* __restrict__ records,
                    int32_t* __restrict__ sites, int skip_big, int skip_small) {
  __shared__ uint64_t lds[SORT_CAP];
  const int tile = blockIdx.x;
  const int64_t n = offsets[tile + 1] - offsets[tile];
  if ((skip_big && n > SORT_CAP) || (skip_small && n <= SMALL_CAP)) return;
  sort_tile<SORT_CAP, SORT_THREADS>(offsets, tile, records, sites, lds);
}

// Most tiles hold a few hundred records: one wave sorts one of them in its
// own LDS slice, with no workgroup barriers (a wave's LDS operations run in
// order, so wave_sync orders the network's steps).
__global__ void __launch_bounds__(SORT_THREADS)
    bin_sort_small_kernel(const int64_t* __restrict__ offsets, int n_tiles, uint64_t* __restrict__ records,
                          int32_t* __restrict__ sites) {
  __shared__ uint64_t lds_all[SORT_THREADS / WAVE][SMALL_CAP];
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int tile = blockIdx.x * (SORT_THREADS / WAVE) + wv;
  if (tile >= n_tiles) return;
  const int64_t a = offsets[tile];
  const int n = (int)(offsets[tile + 1] - a);
  if (n <= 0 || n > SMALL_CAP) return;
  uint64_t* lds = lds_all[wv];
  const uint64_t* g = records + a;
  if (n == 1) {
    if (lane == 0) sites[a] = (int32_t)(uint32_t)g[0];
    return;
  }
  int P = 1;
  while (P < n) P <<= 1;
  const int lgP = ilog2(P);
  for (int i = lane; i < n; i += WAVE) lds[i] = g[i];
  wave_sync();
  for (int lk = 1; lk <= lgP; lk++) {
    bitonic_step<WAVE>(lds, P / 2, n, lk - 1, true, lane);
    wave_sync();
    for (int lj = lk - 2; lj >= 0; lj--) {
      bitonic_step<WAVE>(lds, P / 2, n, lj, false, lane);
      wave_sync();
    }
  }
  for (int i = lane; i < n; i += WAVE) sites[a + i] = (int32_t)(uint32_t)lds[i];
}

// The tiles longer than SORT_CAP (a handful per view, up to ~20k records)
// bound the sort when one 256-thread workgroup merges them through global
// memory: list them, then sort each with 1024 threads and 16k records in
// LDS (128 KiB).
constexpr int BIG_THREADS = 1024;
constexpr int BIG_CAP = 16384;
constexpr int BIG_GRID = 256;

__global__ void __launch_bounds__(1024) big_tiles_kernel(const int64_t* __restrict__ offsets, int n_tiles,
                                                        int32_t* __restrict__ list) {
  __shared__ int n_big;
  if (threadIdx.x == 0) n_big = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < n_tiles; t += blockDim.x)
    if (offsets[t + 1] - offsets[t] > SORT_CAP) list[1 + atomicAdd(&n_big, 1)] = t;
  __syncthreads();
  if (threadIdx.x == 0) list[0] = n_big;
}

__global__ void __launch_bounds__(BIG_THREADS)
    bin_sort_big_kernel(const int64_t* __restrict__ offsets, const int32_t* __restrict__ list,
                        uint64_t* __restrict__ records, int32_t* __restrict__ sites) {
  __shared__ uint64_t lds[BIG_CAP];
  const int n_big = list[0];
  for (int k = blockIdx.x; k < n_big; k += gridDim.x) {
    sort_tile<BIG_CAP, BIG_THREADS>(offsets, list[1 + k], records, sites, lds);
    __syncthreads();
  }
}


// ------------------------------------------------ dispatch schedule
// One workgroup turns the tile offsets into the compositing kernels' dispatch
// order: a block-wide exclusive scan of the per-tile segment counts gives the
// segment table, and counting sorts over length buckets give the
// longest-first orders (members of one bucket land in atomic order, which
// only the schedule sees, never a result). Each thread walks a run of tiles
// three times (count, table and histograms, scatter), so no pass reads
// another thread's output from global memory.
constexpr int SCHED_THREADS = 1024, SCHED_BUCKETS = 1024;

__device__ __forceinline__ int len_bucket(int64_t len) {
  if (len < 512) return (int)len;
  return (int)min((int64_t)(512 + ((len - 512) >> 5)), (int64_t)(SCHED_BUCKETS - 1));
}

// Exclusive scan of one int per thread over the workgroup; *total gets the sum.
__device__ int block_excl_scan(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  int x = v;
#pragma unroll
  for (int m = 1; m < WAVE; m <<= 1) {
    const int y = __shfl_up(x, m, WAVE);
    if (lane >= m) x += y;
  }
  if (lane == WAVE - 1) wsum[w] = x;
  __syncthreads();
  int base = 0, tot = 0;
  for (int k = 0; k < SCHED_THREADS / WAVE; k++) {
    const int s = wsum[k];
    if (k < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

__global__ void __launch_bounds__(SCHED_THREADS)
    tile_schedule_kernel(const int64_t* __restrict__ offsets, int nt, int L, int32_t* tile_order,
                         int32_t* tile_first, int32_t* seg_tile, int32_t* seg_order, int32_t* count) {
  // hs: segments by size bucket; ht: tiles by length bucket
  __shared__ int hs[SCHED_BUCKETS], ht[SCHED_BUCKETS];
  __shared__ int wsum[SCHED_THREADS / WAVE];
  const bool segs = tile_first != nullptr;
  for (int i = threadIdx.x; i < SCHED_BUCKETS; i += SCHED_THREADS) hs[i] = ht[i] = 0;
  // each thread owns a run of tiles; everything below walks only its own run
  const int per = (nt + SCHED_THREADS - 1) / SCHED_THREADS;
  const int t0 = min((int)threadIdx.x * per, nt), t1 = min(t0 + per, nt);
  auto cnt = [&](int t) { return offsets[t + 1] - offsets[t]; };
  auto nseg = [&](int64_t c) { return max(1, (int)((c + L - 1) / L)); };
  // size bucket of segment k of a list of c sites (k = ns - 1 holds the rest)
  auto seg_bucket = [&](int64_t c, int k) {
    return (int)min(min(c - (int64_t)k * L, (int64_t)L), (int64_t)(SCHED_BUCKETS - 1));
  };
  if (!segs) {
    // the order alone: tiles interleaved over the threads (coalesced reads),
    // histogram, descending bucket starts, scatter
    if (!tile_order) return;
    __syncthreads();  // the histogram is zeroed
    for (int t = threadIdx.x; t < nt; t += SCHED_THREADS) atomicAdd(&ht[len_bucket(cnt(t))], 1);
    __syncthreads();
    {
      const int bk = SCHED_BUCKETS - 1 - threadIdx.x;
      int tot;
      const int t_start = block_excl_scan(ht[bk], wsum, &tot);
      ht[bk] = t_start;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nt; t += SCHED_THREADS) tile_order[atomicAdd(&ht[len_bucket(cnt(t))], 1)] = t;
    return;
  }
  int local = 0;
  if (segs)
    for (int t = t0; t < t1; t++) local += nseg(cnt(t));
  __syncthreads();
  int total = 0;
  const int run_first = segs ? block_excl_scan(local, wsum, &total) : 0;
  int first = run_first;
  // table, and the histograms of both orders
  for (int t = t0; t < t1; t++) {
    const int64_t c = cnt(t);
    if (tile_order) atomicAdd(&ht[len_bucket(c)], 1);
    if (segs) {
      const int ns = nseg(c);
      tile_first[t] = first;
      for (int k = 0; k < ns; k++) {
        seg_tile[first + k] = t;
        atomicAdd(&hs[seg_bucket(c, k)], 1);
      }
      first += ns;
    }
  }
  if (segs && threadIdx.x == 0) {
    tile_first[nt] = total;
    *count = total;
  }
  __syncthreads();
  // bucket starts in descending bucket order: thread i owns bucket 1023 - i
  {
    const int bk = SCHED_BUCKETS - 1 - threadIdx.x;
    int tot;
    const int s_start = block_excl_scan(hs[bk], wsum, &tot);
    const int t_start = block_excl_scan(ht[bk], wsum, &tot);
    hs[bk] = s_start;
    ht[bk] = t_start;
  }
  __syncthreads();
  // scatter (members of one bucket land in atomic order)
  first = run_first;
  for (int t = t0; t < t1; t++) {
    const int64_t c = cnt(t);
    if (tile_order) tile_order[atomicAdd(&ht[len_bucket(c)], 1)] = t;
    if (segs) {
      const int ns = nseg(c);
      for (int k = 0; k < ns; k++) seg_order[atomicAdd(&hs[seg_bucket(c, k)], 1)] = first + k;
      first += ns;
    }
  }
}

// The three sort kernels of one call; `list` ([n_tiles + 1] int32 or NULL)
// holds the long tiles for bin_sort_big_kernel. big_tiles_kernel writes its
// count word and entries before the sort reads them, so the list never needs
// clearing between calls.
static int launch_sort(const int64_t* offsets, int32_t n_tiles, uint64_t* records, int32_t* sites, int32_t* list,
                       hipStream_t st) {
  const int per = SORT_THREADS / WAVE;
  bin_sort_small_kernel<<<(n_tiles + per - 1) / per, SORT_THREADS, 0, st>>>(offsets, n_tiles, records, sites);
  bin_sort_kernel<<<n_tiles, SORT_THREADS, 0, st>>>(offsets, n_tiles, records, sites, list ? 1 : 0, 1);
  if (list) {
    big_tiles_kernel<<<1, 1024, 0, st>>>(offsets, n_tiles, list);
    bin_sort_big_kernel<<<BIG_GRID, BIG_THREADS, 0, st>>>(offsets, list, records, sites);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

// Long-tile list of pf_bin_sort_static, owned by the library: one copy per
// GPU (1 MiB + 4 B), held until the process ends. Host code uses only the
// device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) int32_t scratch[PF_SORT_STATIC_MAX_TILES + 1];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static int32_t* g_scratch_ptr[MAX_DEVICES] = {};

// Address of the current device's `scratch`, looked up once per device (no
// allocation; the array is part of the loaded code object).
static int static_scratch(int32_t** out) {
  int dev = -1;
  PF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return set_error(PF_ERR_UNSUPPORTED, "device %d beyond %d devices", dev,
                                                      MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(PF_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = static_cast<int32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PF_OK;
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_bin_count(const pf_camera* cam, const float* spheres, const uint8_t* live, const float* bounds,
                 int32_t n_sites, int32_t* tile_counts, void* stream) {
  if (!cam || !tile_counts || n_sites < 0 || (n_sites > 0 && !spheres))
    return set_error(PF_ERR_ARG, "pf_bin_count: bad arguments");
  if (cam->width < 2 || cam->height < 2) return set_error(PF_ERR_ARG, "pf_bin_count: image must be >= 2x2");
  if (cam->ray_maps) return set_error(PF_ERR_UNSUPPORTED, "pf_bin_count: pinhole cameras only");
  if (reinterpret_cast<uintptr_t>(bounds) % 16) return set_error(PF_ERR_ARG, "pf_bin_count: bounds not 16-byte aligned");
  hipStream_t st = as_stream(stream);
  const BinCam bc = make_bincam(*cam);
  const int64_t nt = (int64_t)tiles_h(*cam) * tiles_w(*cam);
  fill_i32_kernel<<<(unsigned)((nt + 1 + 255) / 256), 256, 0, st>>>(tile_counts, nt + 1, 0);
  if (n_sites > 0)
    bin_count_kernel<<<(n_sites + 255) / 256, 256, 0, st>>>(bc, reinterpret_cast<const float4*>(spheres), live,
                                                           reinterpret_cast<const float4*>(bounds), n_sites,
                                                           tile_counts);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_bin_scan(const int32_t* tile_counts, int32_t n_tiles, int64_t* offsets, void* stream) {
  if (!tile_counts || !offsets || n_tiles < 1) return set_error(PF_ERR_ARG, "pf_bin_scan: bad arguments");
  bin_scan_kernel<<<1, 1024, 0, as_stream(stream)>>>(tile_counts, (int64_t)n_tiles + 1, offsets);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_bin_emit(const pf_camera* cam, const float* spheres, const uint8_t* live, const float* bounds,
                int32_t n_sites, const int64_t* offsets, int32_t n_tiles, uint64_t* records, int32_t* cursor,
                uint8_t* visible, int32_t* dropped, void* stream) {
  if (!cam || !offsets || !cursor || n_sites < 0 || (n_sites > 0 && !spheres))
    return set_error(PF_ERR_ARG, "pf_bin_emit: bad arguments");
  if (cam->ray_maps) return set_error(PF_ERR_UNSUPPORTED, "pf_bin_emit: pinhole cameras only");
  if (reinterpret_cast<uintptr_t>(bounds) % 16) return set_error(PF_ERR_ARG, "pf_bin_emit: bounds not 16-byte aligned");
  if (n_tiles != tiles_h(*cam) * tiles_w(*cam)) return set_error(PF_ERR_ARG, "pf_bin_emit: n_tiles mismatch");
  hipStream_t st = as_stream(stream);
  fill_i32_kernel<<<(unsigned)((n_tiles + 255) / 256), 256, 0, st>>>(cursor, n_tiles, 0);
  if (n_sites > 0)
    bin_emit_kernel<<<(n_sites + 255) / 256, 256, 0, st>>>(make_bincam(*cam), reinterpret_cast<const float4*>(spheres),
                                                          live, reinterpret_cast<const float4*>(bounds), n_sites,
                                                          offsets, records, cursor, visible, dropped);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_live_sites(const pf_camera* cam, const float* spheres, int32_t n_sites, const int32_t* adjacency,
                  const int32_t* adjacency_offsets, int64_t n_edges, const uint16_t* diff, int32_t diff_kind,
                  uint8_t* live, float* bounds, void* stream) {
  if (!cam || n_sites < 0 || n_edges < 0 || (n_sites > 0 && (!spheres || !adjacency_offsets || !live)))
    return set_error(PF_ERR_ARG, "pf_live_sites: bad arguments");
  if (n_edges > 0 && !diff && !adjacency)
    return set_error(PF_ERR_ARG, "pf_live_sites: needs the adjacency or its encoding");
  if (diff && diff_kind != PF_FACES_RADIUS && diff_kind != PF_FACES_POWER)
    return set_error(PF_ERR_ARG, "pf_live_sites: unknown face encoding");
  if (cam->ray_maps) return set_error(PF_ERR_UNSUPPORTED, "pf_live_sites: pinhole cameras only");
  if (reinterpret_cast<uintptr_t>(bounds) % 16)
    return set_error(PF_ERR_ARG, "pf_live_sites: bounds not 16-byte aligned");
  if (n_sites == 0) return PF_OK;
  live_sites_kernel<<<(n_sites + 4 * WAVE - 1) / (4 * WAVE), 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const float4*>(spheres), n_sites, adjacency, adjacency_offsets, n_edges,
      reinterpret_cast<const uint2*>(diff), diff_kind == PF_FACES_POWER ? 1 : 0,
      f3{cam->eye[0], cam->eye[1], cam->eye[2]}, live);
  if (bounds)
    site_bounds_kernel<<<(n_sites + 4 * BOUNDS_SPAN - 1) / (4 * BOUNDS_SPAN), 256, 0, as_stream(stream)>>>(
        reinterpret_cast<const float4*>(spheres), n_sites, adjacency, adjacency_offsets, n_edges,
        reinterpret_cast<const uint2*>(diff), diff_kind == PF_FACES_POWER ? 1 : 0,
        f3{cam->eye[0], cam->eye[1], cam->eye[2]}, live, reinterpret_cast<float4*>(bounds));
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int32_t pf_cone_tree_nodes(int32_t width, int32_t height) {
  ConeTree t;
  if (width < 2 || height < 2 || !make_cone_tree(width, height, t)) return -1;
  return t.total;
}

int pf_bin_cones(const pf_camera* cam, float* cones, void* stream) {
  if (!cam || !cones || !cam->ray_maps) return set_error(PF_ERR_ARG, "pf_bin_cones: needs a camera with ray_maps");
  ConeTree t;
  if (cam->width < 2 || cam->height < 2 || !make_cone_tree(cam->width, cam->height, t))
    return set_error(PF_ERR_ARG, "pf_bin_cones: image must be 5..8192 pixels a side");
  hipStream_t st = as_stream(stream);
  float4* c = reinterpret_cast<float4*>(cones);
  const int nl = t.dh[0] * t.dw[0];
  cone_leaf_kernel<<<(nl + 255) / 256, 256, 0, st>>>(cam->ray_maps, cam->width, cam->height, t.dh[0], t.dw[0], c);
  for (int l = 1; l < t.n_levels; l++) {
    const int np = t.dh[l] * t.dw[l];
    cone_merge_kernel<<<(np + 255) / 256, 256, 0, st>>>(c, t.off[l], t.dh[l], t.dw[l], t.off[l - 1], t.dh[l - 1],
                                                          t.dw[l - 1]);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_bin_count_cones(const pf_camera* cam, const float* cones, const float* spheres, int32_t n_sites,
                       int32_t* tile_counts, void* stream) {
  if (!cam || !cones || !tile_counts || n_sites < 0 || (n_sites > 0 && !spheres))
    return set_error(PF_ERR_ARG, "pf_bin_count_cones: bad arguments");
  ConeTree t;
  if (cam->width < 2 || cam->height < 2 || !make_cone_tree(cam->width, cam->height, t))
    return set_error(PF_ERR_ARG, "pf_bin_count_cones: image must be 5..8192 pixels a side");
  hipStream_t st = as_stream(stream);
  const int64_t nt = (int64_t)tiles_h(*cam) * tiles_w(*cam);
  fill_i32_kernel<<<(unsigned)((nt + 1 + 255) / 256), 256, 0, st>>>(tile_counts, nt + 1, 0);
  if (n_sites > 0)
    bin_count_cones_kernel<<<(n_sites + 255) / 256, 256, 0, st>>>(
        make_bincam(*cam), t, reinterpret_cast<const float4*>(cones), reinterpret_cast<const float4*>(spheres),
        n_sites, tile_counts);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_bin_emit_cones(const pf_camera* cam, const float* cones, const float* spheres, int32_t n_sites,
                      const int64_t* offsets, int32_t n_tiles, uint64_t* records, int32_t* cursor, uint8_t* visible,
                      int32_t* dropped, void* stream) {
  if (!cam || !cones || !offsets || !cursor || n_sites < 0 || (n_sites > 0 && !spheres))
    return set_error(PF_ERR_ARG, "pf_bin_emit_cones: bad arguments");
  ConeTree t;
  if (cam->width < 2 || cam->height < 2 || !make_cone_tree(cam->width, cam->height, t))
    return set_error(PF_ERR_ARG, "pf_bin_emit_cones: image must be 5..8192 pixels a side");
  if (n_tiles != tiles_h(*cam) * tiles_w(*cam)) return set_error(PF_ERR_ARG, "pf_bin_emit_cones: n_tiles mismatch");
  hipStream_t st = as_stream(stream);
  fill_i32_kernel<<<(unsigned)((n_tiles + 255) / 256), 256, 0, st>>>(cursor, n_tiles, 0);
  if (n_sites > 0)
    bin_emit_cones_kernel<<<(n_sites + 255) / 256, 256, 0, st>>>(
        make_bincam(*cam), t, reinterpret_cast<const float4*>(cones), reinterpret_cast<const float4*>(spheres),
        n_sites, offsets, records, cursor, visible, dropped);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

size_t pf_bin_sort_workspace_size(int64_t n_records, int32_t n_tiles) {
  (void)n_records;
  return n_tiles > 0 ? sizeof(int32_t) * ((size_t)n_tiles + 1) : 0;  // the list of long tiles
}

int pf_bin_sort(const int64_t* offsets, int32_t n_tiles, uint64_t* records, int32_t* sites, void* workspace,
                size_t workspace_bytes, void* stream) {
  if (!offsets || n_tiles < 1) return set_error(PF_ERR_ARG, "pf_bin_sort: bad arguments");
  const bool big = workspace && workspace_bytes >= pf_bin_sort_workspace_size(0, n_tiles);
  return launch_sort(offsets, n_tiles, records, sites, big ? static_cast<int32_t*>(workspace) : nullptr,
                     as_stream(stream));
}

int pf_bin_sort_static(const int64_t* offsets, int32_t n_tiles, uint64_t* records, int32_t* sites, void* stream) {
  if (!offsets || n_tiles < 1) return set_error(PF_ERR_ARG, "pf_bin_sort_static: bad arguments");
  if (n_tiles > PF_SORT_STATIC_MAX_TILES)
    return set_error(PF_ERR_OVERFLOW, "pf_bin_sort_static: %d tiles exceed the static scratch of %d tiles", n_tiles,
                     PF_SORT_STATIC_MAX_TILES);
  int32_t* list = nullptr;
  const int s = static_scratch(&list);
  if (s != PF_OK) return s;
  return launch_sort(offsets, n_tiles, records, sites, list, as_stream(stream));
}

int pf_tile_schedule(const int64_t* offsets, int32_t n_tiles, int64_t n_records, int32_t seg_len,
                     int32_t max_segs, int32_t* tile_order, int32_t* tile_first, int32_t* seg_tile,
                     int32_t* seg_order, int32_t* count, void* stream) {
  if (!offsets || n_tiles < 1 || n_records < 0) return set_error(PF_ERR_ARG, "pf_tile_schedule: bad arguments");
  const bool segs = tile_first || seg_tile || seg_order || count;
  if (segs) {
    if (!tile_first || !seg_tile || !seg_order || !count)
      return set_error(PF_ERR_ARG, "pf_tile_schedule: the segment outputs go together");
    if (seg_len < WAVE || seg_len % WAVE || seg_len > SCHED_BUCKETS)
      return set_error(PF_ERR_ARG, "pf_tile_schedule: seg_len must be a multiple of 64 up to 1024");
    if ((int64_t)max_segs < (int64_t)n_tiles + n_records / seg_len)
      return set_error(PF_ERR_ARG, "pf_tile_schedule: max_segs below n_tiles + n_records / seg_len");
  }
  if (!segs && !tile_order) return PF_OK;
  tile_schedule_kernel<<<1, SCHED_THREADS, 0, as_stream(stream)>>>(offsets, n_tiles, seg_len, tile_order,
                                                                    segs ? tile_first : nullptr, seg_tile, seg_order,
                                                                    count);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
