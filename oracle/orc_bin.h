/*
 * ORACLE -- test infrastructure only. Nothing in the product links this.
 *
 * FP32 restatement of the reference's tile binning, in the exact operation
 * order the HIP kernels use (compiled with -ffp-contract=off, correctly
 * rounded sqrt/div), so tile membership and sort keys compare bit-exact.
 *
 * Pinned by (tests/test_oracle_binning.py): coverage -- every pixel whose
 * FP64 ray meets a non-culled ball lies in a listed tile; and a two-sided
 * check of every (site, tile) pair against an ellipse fitted in FP64 to the
 * sampled, projected silhouette of the ball (no shared conic algebra): listed
 * tiles pass its pixel-range and two slab tests within 2e-4, tiles that pass
 * them by more are listed. Planted mistakes (an always-true tile test,
 * semi-axes scaled by 1.02, a shifted centre, swapped axes) fail that check.
 * Parity unpinned: the reference runs under Warp/NVRTC, whose FMA
 * contraction cannot be reproduced; bit-exactness is defined against this
 * mirror. Two Warp semantics are fixed here by choice: wp.cast(float,
 * uint32) is read as a bit reinterpretation, and integer // truncates toward
 * zero (C division).
 */
#ifndef ORC_BIN_H
#define ORC_BIN_H

#include <math.h>
#include <stdint.h>
#include <string.h>
#include <limits.h>

#include "../include/powerfoam_amd.h"

typedef struct { float x, y, z; } v3f;

static inline v3f v3f_make(float x, float y, float z) { v3f r = {x, y, z}; return r; }
static inline v3f v3f_sub(v3f a, v3f b) { return v3f_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline float v3f_dot(v3f a, v3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3f v3f_cross(v3f a, v3f b) {
  return v3f_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float v3f_len(v3f a) { return sqrtf(v3f_dot(a, a)); }
/* wp.normalize: a / |a| when |a| > 0, else 0 */
static inline v3f v3f_normalize(v3f a) {
  float l = v3f_len(a);
  if (l > 0.0f) return v3f_make(a.x / l, a.y / l, a.z / l);
  return v3f_make(0.0f, 0.0f, 0.0f);
}

/* float -> int32 as CUDA/AMD hardware converts: truncate, saturate, NaN -> 0 */
static inline int32_t orc_f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT_MAX;
  if (x <= -2147483648.0f) return INT_MIN;
  return (int32_t)x;
}

typedef struct { float uc, vc, r1, r2, vx, vy; } orc_obb;

/* camera.py:79-128 proj_sphere_to_obb (x**2.0 taken as x*x) */
static inline orc_obb orc_proj_sphere_to_obb(const pf_camera* cam, v3f center, float radius) {
  orc_obb o;
  v3f eye = v3f_make(cam->eye[0], cam->eye[1], cam->eye[2]);
  v3f X = v3f_make(cam->right[0], cam->right[1], cam->right[2]);
  v3f Y = v3f_make(cam->up[0], cam->up[1], cam->up[2]);
  v3f C = v3f_sub(center, eye);
  float C2 = v3f_dot(C, C);
  float R2 = radius * radius;
  v3f Z = v3f_normalize(v3f_cross(Y, X));
  float cx = v3f_dot(C, X), cy = v3f_dot(C, Y), cz = v3f_dot(C, Z);
  float k = C2 - R2;
  float a = cx * cx - k * v3f_dot(X, X);
  float b = cy * cy - k * v3f_dot(Y, Y);
  float f = cz * cz - k * v3f_dot(Z, Z);
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

/* camera.py:131-178 verify_tile_obb_intersection (two OBB axes only) */
static inline int orc_tile_obb_test(int tile_i, int tile_j, const pf_camera* cam, const orc_obb* o) {
  float wm1 = (float)(cam->width - 1), hm1 = (float)(cam->height - 1);
  float u_min = 2.0f * (float)(tile_j * PF_TILE_WIDTH) / wm1 - 1.0f;
  float u_max = 2.0f * (float)((tile_j + 1) * PF_TILE_WIDTH) / wm1 - 1.0f;
  float v_max = 1.0f - 2.0f * (float)(tile_i * PF_TILE_WIDTH) / hm1;
  float v_min = 1.0f - 2.0f * (float)((tile_i + 1) * PF_TILE_WIDTH) / hm1;
  float u_center = (u_min + u_max) * 0.5f;
  float v_center = (v_min + v_max) * 0.5f;
  float eu = (u_max - u_min) * 0.5f;
  float ev = (v_max - v_min) * 0.5f;
  float du = u_center - o->uc;
  float dv = v_center - o->vc;
  float p1 = du * o->vx + dv * o->vy;
  float t1 = eu * fabsf(o->vx) + ev * fabsf(o->vy);
  if (fabsf(p1) > o->r1 + t1) return 0;
  float p2 = du * -o->vy + dv * o->vx;
  float t2 = eu * fabsf(-o->vy) + ev * fabsf(o->vx);
  if (fabsf(p2) > o->r2 + t2) return 0;
  return 1;
}

/* Per-site footprint: cull (rasterize.py:155-160), OBB, pixel range
 * (:162-185) and tile range (:187-190). Returns 0 when culled. */
typedef struct { int ti0, ti1, tj0, tj1; orc_obb obb; uint32_t key; } orc_footprint;

static inline int orc_site_footprint(const pf_camera* cam, const float* sphere, orc_footprint* fp) {
  v3f center = v3f_make(sphere[0], sphere[1], sphere[2]);
  float radius = sphere[3];
  v3f eye = v3f_make(cam->eye[0], cam->eye[1], cam->eye[2]);
  v3f up = v3f_make(cam->up[0], cam->up[1], cam->up[2]);
  v3f right = v3f_make(cam->right[0], cam->right[1], cam->right[2]);
  v3f v = v3f_sub(center, eye);
  v3f fwd = v3f_cross(up, right); /* not normalised, as the reference */
  if (v3f_len(v) < 4.0f * radius || v3f_dot(v, fwd) < 0.1f) return 0;
  orc_obb o = orc_proj_sphere_to_obb(cam, center, radius);
  float ext_u = fabsf(o.vx * o.r1) + fabsf(o.vy * o.r2);
  float ext_v = fabsf(o.vy * o.r1) + fabsf(o.vx * o.r2);
  float u_min = o.uc - ext_u, u_max = o.uc + ext_u;
  float v_min = o.vc - ext_v, v_max = o.vc + ext_v;
  float wm1 = (float)(cam->width - 1), hm1 = (float)(cam->height - 1);
  float j_min_f = (u_min + 1.0f) * 0.5f * wm1;
  float j_max_f = (u_max + 1.0f) * 0.5f * wm1;
  float i_min_f = (1.0f - v_max) * 0.5f * hm1;
  float i_max_f = (1.0f - v_min) * 0.5f * hm1;
  int i_min = orc_f2i(floorf(i_min_f)); if (i_min < 0) i_min = 0;
  int i_max = orc_f2i(ceilf(i_max_f)); if (i_max > cam->height - 1) i_max = cam->height - 1;
  int j_min = orc_f2i(floorf(j_min_f)); if (j_min < 0) j_min = 0;
  int j_max = orc_f2i(ceilf(j_max_f)); if (j_max > cam->width - 1) j_max = cam->width - 1;
  fp->ti0 = i_min / PF_TILE_WIDTH; fp->ti1 = i_max / PF_TILE_WIDTH;
  fp->tj0 = j_min / PF_TILE_WIDTH; fp->tj1 = j_max / PF_TILE_WIDTH;
  fp->obb = o;
  /* rasterize.py:360-363 pow_dist bits */
  float pow_dist = v3f_dot(v, v) - radius * radius;
  uint32_t bits; memcpy(&bits, &pow_dist, 4);
  fp->key = bits;
  return 1;
}

/* IEEE binary16 round-to-nearest-even (the reference's wp.float16(x)). */
static inline uint16_t orc_f32_to_f16(float f) {
  uint32_t x; memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) { /* inf / nan */
    return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0u));
  }
  if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to inf */
  if (absx < 0x38800000u) { /* subnormal half or zero */
    if (absx < 0x33000000u) return (uint16_t)sign; /* < half of min subnormal */
    uint32_t mant = (absx & 0x7fffffu) | 0x800000u;
    int shift = 126 - (int)(absx >> 23); /* 14 .. 24 */
    uint32_t q = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (absx >> 23) - 112u;
  uint32_t m = absx & 0x7fffffu;
  uint32_t h = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}

static inline float orc_f16_to_f32(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
  uint32_t m = (uint32_t)h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else { /* subnormal */
      float v = (float)m * 5.9604644775390625e-08f; /* 2^-24 */
      memcpy(&x, &v, 4);
      x |= sign;
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f; memcpy(&f, &x, 4);
  return f;
}

/* ------------------------------------------------ non-pinhole cone tree
 * _precompute_cones (rasterize.py:1518-1585): level 0 holds one cone per 2x2
 * pixel block (LEAF_SIZE, rasterize.py:12), each further level one per 2x2
 * block of the level below (dims rounded up) until 1x1; nodes are stored
 * level after level, row-major inside a level. Level 2 has exactly the 8x8
 * tiles (tile_level, :1546-1550). A cone is (axis, cos half-angle), cos > 1
 * marking a node without valid rays. */
#define ORC_LEAF_SIZE 2
#define ORC_MAX_LEVELS 32

typedef struct {
  int n_levels, tile_level, total;
  int off[ORC_MAX_LEVELS], dh[ORC_MAX_LEVELS], dw[ORC_MAX_LEVELS];
} orc_cone_tree;

static inline void orc_cone_tree_dims(int width, int height, orc_cone_tree* t) {
  int dh = (height + ORC_LEAF_SIZE - 1) / ORC_LEAF_SIZE, dw = (width + ORC_LEAF_SIZE - 1) / ORC_LEAF_SIZE;
  t->n_levels = 0;
  t->total = 0;
  while (t->n_levels < ORC_MAX_LEVELS) {
    t->off[t->n_levels] = t->total;
    t->dh[t->n_levels] = dh;
    t->dw[t->n_levels] = dw;
    t->total += dh * dw;
    t->n_levels++;
    if (dh == 1 && dw == 1) break;
    dh = (dh + 1) / 2;
    dw = (dw + 1) / 2;
  }
  t->tile_level = 0;
  for (int sz = ORC_LEAF_SIZE; sz < PF_TILE_WIDTH; sz *= 2) t->tile_level++;
}

/* compute_leaf_cones_kernel (rasterize.py:468-512) and merge_cones_kernel
 * (:514-569), level by level; cones is [total, 4]. */
static inline void orc_build_cones(const pf_camera* cam, const orc_cone_tree* t, float* cones) {
  const int W = cam->width, H = cam->height;
  for (int a = 0; a < t->dh[0]; a++)
    for (int b = 0; b < t->dw[0]; b++) {
      float* out = cones + 4 * (size_t)(a * t->dw[0] + b);
      v3f rays[4];
      v3f sum = v3f_make(0.0f, 0.0f, 0.0f);
      int nv = 0;
      for (int di = 0; di < ORC_LEAF_SIZE; di++)
        for (int dj = 0; dj < ORC_LEAF_SIZE; dj++) {
          int pi = a * ORC_LEAF_SIZE + di, pj = b * ORC_LEAF_SIZE + dj;
          if (pi > H - 1) pi = H - 1;
          if (pj > W - 1) pj = W - 1;
          const float* rm = cam->ray_maps + ((size_t)pi * W + pj) * 6;
          v3f r = v3f_make(rm[3], rm[4], rm[5]);
          float len = v3f_len(r);
          if (len > 1.0e-6f) {
            v3f rn = v3f_make(r.x / len, r.y / len, r.z / len);
            rays[nv++] = rn;
            sum = v3f_make(sum.x + rn.x, sum.y + rn.y, sum.z + rn.z);
          }
        }
      if (nv == 0) {
        out[0] = out[1] = out[2] = 0.0f;
        out[3] = 2.0f;
        continue;
      }
      v3f axis = v3f_normalize(sum);
      float ch = 1.0f;
      for (int k = 0; k < nv; k++) {
        float d = v3f_dot(axis, rays[k]);
        ch = d < ch ? d : ch;
      }
      out[0] = axis.x; out[1] = axis.y; out[2] = axis.z; out[3] = ch;
    }
  for (int l = 1; l < t->n_levels; l++) {
    const int pdw = t->dw[l], cdh = t->dh[l - 1], cdw = t->dw[l - 1];
    for (int a = 0; a < t->dh[l]; a++)
      for (int b = 0; b < pdw; b++) {
        float* out = cones + 4 * (size_t)(t->off[l] + a * pdw + b);
        v3f ax[4];
        float cc[4];
        v3f sum = v3f_make(0.0f, 0.0f, 0.0f);
        int nv = 0;
        for (int di = 0; di < 2; di++)
          for (int dj = 0; dj < 2; dj++) {
            int ca = 2 * a + di, cb = 2 * b + dj;
            if (ca < cdh && cb < cdw) {
              const float* c = cones + 4 * (size_t)(t->off[l - 1] + ca * cdw + cb);
              if (c[3] <= 1.0f) {
                ax[nv] = v3f_make(c[0], c[1], c[2]);
                cc[nv] = c[3];
                sum = v3f_make(sum.x + c[0], sum.y + c[1], sum.z + c[2]);
                nv++;
              }
            }
          }
        if (nv == 0) {
          out[0] = out[1] = out[2] = 0.0f;
          out[3] = 2.0f;
          continue;
        }
        v3f axis = v3f_normalize(sum);
        float ch = 1.0f;
        for (int k = 0; k < nv; k++) {
          float d = v3f_dot(axis, ax[k]);
          float sa = 1.0f - d * d;
          float sc = 1.0f - cc[k] * cc[k];
          float sin_angle = sqrtf(sa > 0.0f ? sa : 0.0f);
          float sin_child = sqrtf(sc > 0.0f ? sc : 0.0f);
          float ext = d * cc[k] - sin_angle * sin_child;
          ch = ext < ch ? ext : ch;
        }
        out[0] = axis.x; out[1] = axis.y; out[2] = axis.z; out[3] = ch;
      }
  }
}

/* camera.py:27-44 sphere_cone_intersect */
static inline int orc_sphere_cone(v3f center, float radius, v3f eye, const float* cone) {
  v3f v = v3f_sub(center, eye);
  float dist = v3f_len(v);
  float cos_alpha = v3f_dot(v, v3f_make(cone[0], cone[1], cone[2])) / dist;
  float sin_half = sqrtf(1.0f - cone[3] * cone[3]);
  float sin_beta = radius / dist;
  float cos_beta = sqrtf(1.0f - sin_beta * sin_beta);
  float threshold = cone[3] * cos_beta - sin_half * sin_beta;
  return cos_alpha > threshold;
}

/* count_visible_generic_kernel / write_visible_generic_kernel
 * (rasterize.py:202-291, 365-466): the cull of the pinhole path, then a
 * depth-first descent (32-entry stack, children pushed in (0,0), (0,1),
 * (1,0), (1,1) order) through the nodes whose cone meets the sphere; tiles
 * reached at tile_level are written to tiles_out in visiting order. Returns
 * the tile count, or -1 when the site is culled; *key gets the depth key. */
static inline int orc_site_cone_tiles(const pf_camera* cam, const orc_cone_tree* t, const float* cones,
                                      const float* sphere, int32_t* tiles_out, uint32_t* key) {
  v3f center = v3f_make(sphere[0], sphere[1], sphere[2]);
  float radius = sphere[3];
  v3f eye = v3f_make(cam->eye[0], cam->eye[1], cam->eye[2]);
  v3f up = v3f_make(cam->up[0], cam->up[1], cam->up[2]);
  v3f right = v3f_make(cam->right[0], cam->right[1], cam->right[2]);
  v3f v = v3f_sub(center, eye);
  v3f fwd = v3f_cross(up, right);
  if (v3f_len(v) < 4.0f * radius || v3f_dot(v, fwd) < 0.1f) return -1;
  float pow_dist = v3f_dot(v, v) - radius * radius;
  memcpy(key, &pow_dist, 4);
  const int th = 1 + (cam->height - 1) / PF_TILE_WIDTH;
  int sl[32], sa[32], sb[32];
  int n = 0, sz = 1;
  sl[0] = t->n_levels - 1;
  sa[0] = 0;
  sb[0] = 0;
  while (sz > 0) {
    sz--;
    const int level = sl[sz], a = sa[sz], b = sb[sz];
    const float* cone = cones + 4 * (size_t)(t->off[level] + a * t->dw[level] + b);
    if (cone[3] > 1.0f) continue;
    if (!orc_sphere_cone(center, radius, eye, cone)) continue;
    if (level == t->tile_level) {
      tiles_out[n++] = a + b * th;
      continue;
    }
    const int cl = level - 1;
    for (int di = 0; di < 2; di++)
      for (int dj = 0; dj < 2; dj++) {
        int ca = 2 * a + di, cb = 2 * b + dj;
        if (ca < t->dh[cl] && cb < t->dw[cl]) {
          sl[sz] = cl;
          sa[sz] = ca;
          sb[sz] = cb;
          sz++;
        }
      }
  }
  return n;
}

#endif /* ORC_BIN_H */
