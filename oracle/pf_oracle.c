/*
 * ORACLE -- test infrastructure only: the CPU restatement of the reference
 * (theialab/powerfoam) renderer hot path that tests/, the smoke check and the
 * bench's CPU baseline use as the checker. The product never links or loads
 * it. Built by oracle/Makefile into oracle/build/liborcl.so.
 *
 *   orc_bin_count / orc_bin_lists   rasterize.py:140-200, 293-363, 1713-1768;
 *                                   non-pinhole: 202-291, 365-569, 1518-1585
 *   orc_adj_encode                  rasterize.py:580-606
 *   orc_raster_fwd_{f32,f64}        rasterize.py:892-1146
 *   orc_raster_bwd_{f32,f64}        rasterize.py:1150-1514
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc_bin.h"

/* ---------------------------------------------------------------- binning */

static int orc_tiles_h(const pf_camera* cam) { return 1 + (cam->height - 1) / PF_TILE_WIDTH; }
static int orc_tiles_w(const pf_camera* cam) { return 1 + (cam->width - 1) / PF_TILE_WIDTH; }

/* Non-pinhole cameras (cam->ray_maps set) bin through the cone tree. */
typedef struct {
  orc_cone_tree t;
  float* cones;
  int32_t* tiles;
} orc_cone_ctx;

static void orc_cone_ctx_init(const pf_camera* cam, orc_cone_ctx* c) {
  orc_cone_tree_dims(cam->width, cam->height, &c->t);
  c->cones = (float*)malloc(sizeof(float) * 4 * (size_t)c->t.total);
  c->tiles = (int32_t*)malloc(sizeof(int32_t) * (size_t)(orc_tiles_h(cam) * orc_tiles_w(cam)));
  orc_build_cones(cam, &c->t, c->cones);
}

static void orc_cone_ctx_free(orc_cone_ctx* c) {
  free(c->cones);
  free(c->tiles);
}

/* Node count, level count and tile level of the cone tree of a W x H image. */
void orc_cone_tree_info(int width, int height, int32_t* out3) {
  orc_cone_tree t;
  orc_cone_tree_dims(width, height, &t);
  out3[0] = t.total;
  out3[1] = t.n_levels;
  out3[2] = t.tile_level;
}

/* The cone tree of cam (ray_maps required), [nodes, 4]. */
void orc_cones(const pf_camera* cam, float* cones) {
  orc_cone_tree t;
  orc_cone_tree_dims(cam->width, cam->height, &t);
  orc_build_cones(cam, &t, cones);
}

/* Tiles one site is binned to through the cone tree (visiting order); -1 if culled. */
int orc_site_cone_tiles_dbg(const pf_camera* cam, const float* sphere, int32_t* tiles_out) {
  orc_cone_ctx c;
  orc_cone_ctx_init(cam, &c);
  uint32_t key;
  int n = orc_site_cone_tiles(cam, &c.t, c.cones, sphere, tiles_out, &key);
  orc_cone_ctx_free(&c);
  return n;
}

/* counts[t+1] = number of sites binned to tile t; returns the total. */
int64_t orc_bin_count(const pf_camera* cam, const float* spheres, int32_t n, int32_t* counts) {
  const int th = orc_tiles_h(cam), tw = orc_tiles_w(cam);
  memset(counts, 0, sizeof(int32_t) * (size_t)(th * tw + 1));
  int64_t total = 0;
  if (cam->ray_maps) {
    orc_cone_ctx c;
    orc_cone_ctx_init(cam, &c);
    for (int s = 0; s < n; s++) {
      uint32_t key;
      int k = orc_site_cone_tiles(cam, &c.t, c.cones, spheres + 4 * (size_t)s, c.tiles, &key);
      for (int i = 0; i < k; i++) counts[c.tiles[i] + 1]++;
      total += k > 0 ? k : 0;
    }
    orc_cone_ctx_free(&c);
    return total;
  }
  for (int s = 0; s < n; s++) {
    orc_footprint fp;
    if (!orc_site_footprint(cam, spheres + 4 * (size_t)s, &fp)) continue;
    for (int ti = fp.ti0; ti <= fp.ti1; ti++)
      for (int tj = fp.tj0; tj <= fp.tj1; tj++) {
        if (!orc_tile_obb_test(ti, tj, cam, &fp.obb)) continue;
        counts[ti + tj * th + 1]++;
        total++;
      }
  }
  return total;
}

typedef struct { uint32_t key; int32_t site; } orc_rec;

static int orc_rec_cmp(const void* a, const void* b) {
  const orc_rec* x = (const orc_rec*)a;
  const orc_rec* y = (const orc_rec*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return (x->site > y->site) - (x->site < y->site);
}

/* Fills the per-tile lists given offsets (cumsum of counts), each tile's
 * entries ordered by (depth key, site). keys_out may be NULL. */
void orc_bin_lists(const pf_camera* cam, const float* spheres, int32_t n, const int64_t* offsets,
                   int32_t* sites_out, uint32_t* keys_out) {
  const int th = orc_tiles_h(cam), tw = orc_tiles_w(cam);
  const int nt = th * tw;
  int64_t m = offsets[nt];
  orc_rec* recs = (orc_rec*)malloc(sizeof(orc_rec) * (size_t)(m > 0 ? m : 1));
  int64_t* cur = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
  if (cam->ray_maps) {
    orc_cone_ctx c;
    orc_cone_ctx_init(cam, &c);
    for (int s = 0; s < n; s++) {
      uint32_t key;
      int k = orc_site_cone_tiles(cam, &c.t, c.cones, spheres + 4 * (size_t)s, c.tiles, &key);
      for (int i = 0; i < k; i++) {
        int t = c.tiles[i];
        int64_t idx = offsets[t] + cur[t]++;
        if (idx < offsets[t + 1]) {
          recs[idx].key = key;
          recs[idx].site = s;
        }
      }
    }
    orc_cone_ctx_free(&c);
  }
  for (int s = 0; s < n && !cam->ray_maps; s++) {
    orc_footprint fp;
    if (!orc_site_footprint(cam, spheres + 4 * (size_t)s, &fp)) continue;
    for (int ti = fp.ti0; ti <= fp.ti1; ti++)
      for (int tj = fp.tj0; tj <= fp.tj1; tj++) {
        if (!orc_tile_obb_test(ti, tj, cam, &fp.obb)) continue;
        int t = ti + tj * th;
        int64_t idx = offsets[t] + cur[t]++;
        if (idx < offsets[t + 1]) {
          recs[idx].key = fp.key;
          recs[idx].site = s;
        }
      }
  }
  for (int t = 0; t < nt; t++) {
    int64_t a = offsets[t], b = offsets[t + 1];
    if (b - a > 1) qsort(recs + a, (size_t)(b - a), sizeof(orc_rec), orc_rec_cmp);
  }
  for (int64_t i = 0; i < m; i++) {
    sites_out[i] = recs[i].site;
    if (keys_out) keys_out[i] = recs[i].key;
  }
  free(cur);
  free(recs);
}

/* Exposes the per-site footprint for the brute-force tests:
 * out[0..3] = tile range (ti0, ti1, tj0, tj1), obb[0..5]; returns 0 if culled. */
int orc_site_footprint_dbg(const pf_camera* cam, const float* sphere, int32_t* range, float* obb) {
  orc_footprint fp;
  if (!orc_site_footprint(cam, sphere, &fp)) return 0;
  range[0] = fp.ti0; range[1] = fp.ti1; range[2] = fp.tj0; range[3] = fp.tj1;
  obb[0] = fp.obb.uc; obb[1] = fp.obb.vc; obb[2] = fp.obb.r1;
  obb[3] = fp.obb.r2; obb[4] = fp.obb.vx; obb[5] = fp.obb.vy;
  return 1;
}

int orc_tile_obb_test_dbg(int tile_i, int tile_j, const pf_camera* cam, const float* obb) {
  orc_obb o = {obb[0], obb[1], obb[2], obb[3], obb[4], obb[5]};
  return orc_tile_obb_test(tile_i, tile_j, cam, &o);
}

/* ------------------------------------------------------- adjacency encode */

void orc_adj_encode(const float* spheres, int32_t n, const int32_t* adj, const int32_t* adj_off,
                    uint16_t* diff) {
  for (int s = 0; s < n; s++) {
    const float* sp = spheres + 4 * (size_t)s;
    for (int e = adj_off[s]; e < adj_off[s + 1]; e++) {
      const float* ap = spheres + 4 * (size_t)adj[e];
      diff[4 * (size_t)e + 0] = orc_f32_to_f16(ap[0] - sp[0]);
      diff[4 * (size_t)e + 1] = orc_f32_to_f16(ap[1] - sp[1]);
      diff[4 * (size_t)e + 2] = orc_f32_to_f16(ap[2] - sp[2]);
      diff[4 * (size_t)e + 3] = orc_f32_to_f16(ap[3]);
    }
  }
}

void orc_f32_to_f16_array(const float* x, int64_t n, uint16_t* y) {
  for (int64_t i = 0; i < n; i++) y[i] = orc_f32_to_f16(x[i]);
}

/* benchmark.py:297-302: (c_adj - c, pm_adj - pm), pm = (|c|^2 - r^2)/2, in
 * binary16; the differences are formed in float32 as torch does. */
void orc_power_face_encode(const float* spheres, int32_t n, const int32_t* adj, const int32_t* adj_off,
                           uint16_t* diff) {
  for (int s = 0; s < n; s++) {
    const float* sp = spheres + 4 * (size_t)s;
    float pm_s = 0.5f * ((sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]) - sp[3] * sp[3]);
    for (int e = adj_off[s]; e < adj_off[s + 1]; e++) {
      const float* ap = spheres + 4 * (size_t)adj[e];
      float pm_a = 0.5f * ((ap[0] * ap[0] + ap[1] * ap[1] + ap[2] * ap[2]) - ap[3] * ap[3]);
      diff[4 * (size_t)e + 0] = orc_f32_to_f16(ap[0] - sp[0]);
      diff[4 * (size_t)e + 1] = orc_f32_to_f16(ap[1] - sp[1]);
      diff[4 * (size_t)e + 2] = orc_f32_to_f16(ap[2] - sp[2]);
      diff[4 * (size_t)e + 3] = orc_f32_to_f16(pm_a - pm_s);
    }
  }
}

void orc_f16_to_f32_array(const uint16_t* x, int64_t n, float* y) {
  for (int64_t i = 0; i < n; i++) y[i] = orc_f16_to_f32(x[i]);
}

/* ------------------------------------------------------------- compositing */

/* Ill-conditioning test of the backward's `fragile` output (a site some ray
 * meets where its gradient is not resolved): |h^2| < h2_rel r^2 +
 * h2_abs r |o - c| (h^2 = r^2 - |(o - c) x d|^2, a tangent ray; the second
 * term is the rounding a float32 ray carries: a direction correct to a few
 * ulp misplaces the ray by ~eps |o - c| near the site), or a clipped interval
 * within dt_rel (1 + t_far) of zero length. */
static double g_frag_h2_rel = 1e-6, g_frag_h2_abs = 0.0, g_frag_dt_rel = 1e-5;

void orc_set_fragile(double h2_rel, double h2_abs, double dt_rel) {
  g_frag_h2_rel = h2_rel;
  g_frag_h2_abs = h2_abs;
  g_frag_dt_rel = dt_rel;
}

#define R float
#define SFX _f32
#include "orc_raster_impl.h"
#undef R
#undef SFX

#define R double
#define SFX _f64
#include "orc_raster_impl.h"
#undef R
#undef SFX

/* ------------------------------------- restricted power adjacency (FP64) */
#include "orc_power.h"

/* ------------------------------------------------ ABI layout (for tests) */
#include <stddef.h>
/* sizeof and field offsets of the C-ABI structs as gcc lays them out. */
int orc_abi_layout(int64_t* out, int cap) {
  int64_t v[] = {
      (int64_t)sizeof(pf_camera), (int64_t)offsetof(pf_camera, ray_maps), (int64_t)offsetof(pf_camera, fov_cos_cutoff),
      (int64_t)sizeof(pf_scene), (int64_t)offsetof(pf_scene, adjacency_diff),
      (int64_t)offsetof(pf_scene, adjacency_count),
      (int64_t)sizeof(pf_tile_lists), (int64_t)offsetof(pf_tile_lists, sites),
      (int64_t)sizeof(pf_raster_fwd_args), (int64_t)offsetof(pf_raster_fwd_args, tiles),
      (int64_t)offsetof(pf_raster_fwd_args, transmittance_threshold), (int64_t)offsetof(pf_raster_fwd_args, point_err),
      (int64_t)offsetof(pf_raster_fwd_args, entry_shades), (int64_t)offsetof(pf_raster_fwd_args, mlp_head),
      (int64_t)sizeof(pf_mlp_head), (int64_t)offsetof(pf_mlp_head, latents),
      (int64_t)offsetof(pf_mlp_head, texel_rgbh_out),
      (int64_t)sizeof(pf_raster_bwd_args), (int64_t)offsetof(pf_raster_bwd_args, exact_gradients),
      (int64_t)offsetof(pf_raster_bwd_args, grad_texel_height), (int64_t)offsetof(pf_raster_bwd_args, entry_shades),
      (int64_t)sizeof(pf_raster_rgb_args), (int64_t)offsetof(pf_raster_rgb_args, color),
      (int64_t)sizeof(pf_raster_vis_args), (int64_t)offsetof(pf_raster_vis_args, bkgd_color),
      (int64_t)offsetof(pf_raster_vis_args, intersections),
      (int64_t)sizeof(pf_trace_args), (int64_t)offsetof(pf_trace_args, max_steps),
      (int64_t)offsetof(pf_trace_args, n_cells), (int64_t)offsetof(pf_trace_args, tile_steps),
      (int64_t)offsetof(pf_trace_args, face_planes),
      (int64_t)sizeof(pf_trace_bwd_args), (int64_t)offsetof(pf_trace_bwd_args, grad_opacity),
      (int64_t)offsetof(pf_trace_bwd_args, grad_texel_height), (int64_t)offsetof(pf_trace_bwd_args, tile_order),
      (int64_t)sizeof(pf_mlp_rows), (int64_t)offsetof(pf_mlp_rows, eye), (int64_t)offsetof(pf_mlp_rows, latents),
      (int64_t)offsetof(pf_mlp_rows, site_count), (int64_t)offsetof(pf_mlp_rows, out_stride),
      (int64_t)sizeof(pf_mlp_grads),
      (int64_t)sizeof(pf_segments), (int64_t)offsetof(pf_segments, count), (int64_t)offsetof(pf_segments, seg_len),
      (int64_t)offsetof(pf_segments, scratch)};
  int n = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < n && i < cap; i++) out[i] = v[i];
  return n;
}
