/*
 * ORACLE -- test infrastructure only. Included twice by pf_oracle.c with
 *   R   = float  / double   (arithmetic type)
 *   SFX = _f32   / _f64     (function-name suffix)
 *
 * A CPU restatement of the reference's per-tile compositing, lane by lane,
 * with the two tile-wide decisions the reference makes (the tile-max
 * transmittance early stop, rasterize.py:985-991, and the any-hit skip,
 * :1028-1032) taken over the 64 lanes of the tile.
 *
 *   ray_sphere_intersect        rendering_math.py:4-23
 *   ray_sphere_intersect_bwd    rendering_math.py:26-59
 *   ray_pface_intersect         rendering_math.py:62-82
 *   ray_pface_intersect_bwd     rendering_math.py:101-125
 *   ray_plane_intersect(_bwd)   rendering_math.py:128-147
 *   plane_intersection_fwd      texture.py:10-71
 *   plane_intersection_bwd      texture.py:74-186
 *   forward_kernel              rasterize.py:892-1146
 *   backward_kernel             rasterize.py:1150-1514
 *   get_ray_dir                 camera.py:19-24
 *
 * Pinned by: closed-form chord lengths for one- and two-site scenes, the
 * opacity invariant, the texel special cases, and FP64 central finite
 * differences of the exact-gradient mode (tests/test_oracle_raster.py).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)
#define V3 CAT(v3, SFX)

typedef struct { R x, y, z; } V3;

static inline V3 FN(mk)(R x, R y, R z) { V3 r = {x, y, z}; return r; }
static inline V3 FN(add)(V3 a, V3 b) { return FN(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 FN(sub)(V3 a, V3 b) { return FN(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 FN(scl)(V3 a, R s) { return FN(mk)(a.x * s, a.y * s, a.z * s); }
static inline V3 FN(dvs)(V3 a, R s) { return FN(mk)(a.x / s, a.y / s, a.z / s); }
static inline R FN(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 FN(crs)(V3 a, V3 b) {
  return FN(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline R FN(len)(V3 a) { return (R)sqrt((double)FN(dot)(a, a)); }
static inline V3 FN(nrm)(V3 a) {
  R l = FN(len)(a);
  return l > (R)0 ? FN(dvs)(a, l) : FN(mk)(0, 0, 0);
}
static inline R FN(rmax)(R a, R b) { return a > b ? a : b; }
static inline R FN(rmin)(R a, R b) { return a < b ? a : b; }
static inline R FN(rexp)(R x) { return (R)exp((double)x); }
static inline R FN(rlog)(R x) { return (R)log((double)x); }
static inline R FN(rsqrt_)(R x) { return (R)sqrt((double)x); }
static inline V3 FN(ld3)(const float* p) { return FN(mk)((R)p[0], (R)p[1], (R)p[2]); }

typedef struct {
  int hit;
  R t_near, t_far;
} FN(sphere_hit);

static inline FN(sphere_hit) FN(ray_sphere)(V3 o, V3 d, V3 c, R r) {
  FN(sphere_hit) h = {0, 0, 0};
  V3 oc = FN(sub)(o, c);
  R qb = (R)2 * FN(dot)(oc, d);
  R qc = FN(dot)(oc, oc) - r * r;
  R disc = qb * qb - (R)4 * qc;
  if (disc < (R)0) return h;
  R sq = FN(rsqrt_)(disc);
  R tf = (-qb + sq) / (R)2;
  R tn = (-qb - sq) / (R)2;
  if (tn < (R)0 && tf < (R)0) return h;
  h.hit = 1;
  h.t_near = tn < (R)0 ? (R)0 : tn;
  h.t_far = tf;
  return h;
}

typedef struct { V3 dtn_dc, dtf_dc; R dtn_dr, dtf_dr; } FN(sphere_grad);

static inline FN(sphere_grad) FN(ray_sphere_bwd)(V3 o, V3 d, V3 c, R r) {
  FN(sphere_grad) g;
  g.dtn_dc = FN(mk)(0, 0, 0); g.dtf_dc = FN(mk)(0, 0, 0); g.dtn_dr = 0; g.dtf_dr = 0;
  V3 oc = FN(sub)(o, c);
  R qa = FN(dot)(d, d);
  R qb = (R)2 * FN(dot)(oc, d);
  R qc = FN(dot)(oc, oc) - r * r;
  R disc = qb * qb - (R)4 * qa * qc;
  if (disc < (R)0) return g;
  R sq = FN(rsqrt_)(disc);
  R tf = (-qb + sq) / ((R)2 * qa);
  R tn = (-qb - sq) / ((R)2 * qa);
  if (tn < (R)0 && tf < (R)0) return g;
  /* (2 qa oc - qb d) / (qa sqrt(disc)) */
  V3 w = FN(dvs)(FN(sub)(FN(scl)(oc, (R)2 * qa), FN(scl)(d, qb)), qa * sq);
  V3 dq = FN(dvs)(d, qa);
  g.dtf_dc = FN(add)(dq, w);
  g.dtf_dr = (R)2 * r / sq;
  if (!(tn < (R)0)) {
    g.dtn_dc = FN(sub)(dq, w);
    g.dtn_dr = -(R)2 * r / sq;
  }
  return g;
}

/* face between site (c, r) and neighbour (a, ra) */
static inline void FN(ray_pface)(V3 o, V3 d, V3 c, R r, V3 a, R ra, R* t, R* dp) {
  V3 n = FN(sub)(a, c);
  R off = (R)0.5 * (FN(dot)(a, a) - FN(dot)(c, c) + r * r - ra * ra);
  *dp = FN(dot)(d, n);
  *t = (off - FN(dot)(o, n)) / *dp;
}

static inline void FN(ray_pface_bwd)(V3 o, V3 d, V3 c, R r, V3 a, R ra, V3* dtdc, R* dtdr,
                                     V3* dtda, R* dtdra) {
  V3 n = FN(sub)(a, c);
  R off = (R)0.5 * (FN(dot)(a, a) - FN(dot)(c, c) + r * r - ra * ra);
  R num = off - FN(dot)(o, n);
  R dp = FN(dot)(d, n);
  V3 q = FN(dvs)(FN(dvs)(FN(scl)(d, num), dp), dp);
  *dtdc = FN(add)(FN(dvs)(FN(sub)(o, c), dp), q);
  *dtda = FN(sub)(FN(dvs)(FN(sub)(a, o), dp), q);
  *dtdr = r / dp;
  *dtdra = -ra / dp;
}

static inline void FN(ray_plane)(V3 o, V3 d, V3 p, V3 n, R h, R* t, R* dp) {
  *dp = FN(dot)(n, d);
  *t = (FN(dot)(FN(sub)(p, o), n) + h) / *dp;
}

/* eps = 1e-6 reproduces the reference; eps = 0 is the exact derivative */
static inline void FN(ray_plane_bwd)(V3 o, V3 d, V3 p, V3 n, R h, R eps, V3* dtdp, V3* dtdn,
                                     R* dtdh) {
  R dp = FN(dot)(n, d);
  *dtdp = FN(dvs)(n, dp + eps);
  *dtdh = (R)1 / (dp + eps);
  V3 pe = FN(sub)(p, o);
  *dtdn = FN(sub)(FN(dvs)(pe, dp + eps), FN(dvs)(FN(scl)(d, FN(dot)(pe, n) + h), dp * dp + eps));
}

typedef struct {
  R t_surf0, dp0, height, wsum0, t_surf, dp, wsum;
  V3 rgb;
} FN(plane_fwd_out);

static inline FN(plane_fwd_out) FN(plane_fwd)(V3 o, V3 d, R t_near, V3 p, V3 n, R radius, int hit,
                                              const float* sites, const float* rgbh, int T) {
  FN(plane_fwd_out) r;
  memset(&r, 0, sizeof(r));
  if (!hit) return r;
  R t0, dp0;
  FN(ray_plane)(o, d, p, n, (R)0, &t0, &dp0);
  R tq0 = dp0 >= (R)0 ? t_near : FN(rmax)(t_near, t0);
  V3 x0 = FN(add)(o, FN(scl)(d, tq0));
  R hsum = 0, ws0 = 0;
  for (int i = 0; i < T; i++) {
    V3 s = FN(ld3)(sites + 3 * i);
    V3 df = FN(sub)(x0, s);
    R dist_sq = FN(dot)(df, df) / (radius * radius);
    R w = FN(rexp)(-(R)10 * dist_sq);
    hsum += w * (R)rgbh[4 * i + 3];
    ws0 += w;
  }
  ws0 = FN(rmax)(ws0, (R)1e-20);
  R height = hsum / ws0;
  R t1, dp1;
  FN(ray_plane)(o, d, p, n, height, &t1, &dp1);
  R tq = dp1 >= (R)0 ? t_near : FN(rmax)(t_near, t1);
  V3 x = FN(add)(o, FN(scl)(d, tq));
  V3 rs = FN(mk)(0, 0, 0);
  R ws = 0;
  for (int i = 0; i < T; i++) {
    V3 s = FN(ld3)(sites + 3 * i);
    V3 df = FN(sub)(x, s);
    R dist_sq = FN(dot)(df, df) / (radius * radius);
    R w = FN(rexp)(-(R)10 * dist_sq);
    rs = FN(add)(rs, FN(scl)(FN(ld3)(rgbh + 4 * i), w));
    ws += w;
  }
  ws = FN(rmax)(ws, (R)1e-20);
  r.t_surf0 = t0; r.dp0 = dp0; r.height = height; r.wsum0 = ws0;
  r.t_surf = t1; r.dp = dp1; r.wsum = ws;
  r.rgb = FN(dvs)(rs, ws);
  return r;
}

static inline void FN(atomic_addr)(R* p, R v) {
#pragma omp atomic
  *p += v;
}

/* texture.py:74-186. Gradients of texel sites/rgb/height are accumulated
 * into the given rows; returns dL/dplane_origin, dL/dplane_normal and the
 * gradient w.r.t. the query t_near. */
static inline void FN(plane_bwd)(V3 o, V3 d, R t_near0, V3 p, V3 n, R radius, int hit,
                                 const float* sites, const float* rgbh, const FN(plane_fwd_out) * f,
                                 V3 dLdrgb, R dLdt_surf_in, int T, int exact, R* g_sites, R* g_rgb,
                                 R* g_height, V3* dLdp, V3* dLdn, R* dLdtn) {
  *dLdp = FN(mk)(0, 0, 0);
  *dLdn = FN(mk)(0, 0, 0);
  *dLdtn = 0;
  if (!hit) return;
  const R cut = exact ? (R)0 : (R)1e-3;
  const R eps = exact ? (R)0 : (R)1e-6;
  R r2 = radius * radius;
  R dLdheight = 0;
  /* colour stage */
  R tq = f->dp >= (R)0 ? t_near0 : FN(rmax)(t_near0, f->t_surf);
  V3 x = FN(add)(o, FN(scl)(d, tq));
  V3 dLdx = FN(mk)(0, 0, 0);
  R gdp_in = FN(dot)(dLdrgb, f->rgb);
  for (int i = 0; i < T; i++) {
    V3 s = FN(ld3)(sites + 3 * i);
    V3 df = FN(sub)(x, s);
    R w = FN(rexp)(-(R)10 * (FN(dot)(df, df) / r2));
    if (w / f->wsum < cut) continue;
    V3 dwds = FN(dvs)(FN(scl)(df, (R)2 * (R)10 * w), r2);
    V3 rg = FN(dvs)(FN(scl)(dLdrgb, w), f->wsum);
    FN(atomic_addr)(&g_rgb[3 * i + 0], rg.x);
    FN(atomic_addr)(&g_rgb[3 * i + 1], rg.y);
    FN(atomic_addr)(&g_rgb[3 * i + 2], rg.z);
    R gdp = FN(dot)(dLdrgb, FN(ld3)(rgbh + 4 * i));
    V3 sg = FN(dvs)(FN(scl)(dwds, gdp - gdp_in), f->wsum);
    FN(atomic_addr)(&g_sites[3 * i + 0], sg.x);
    FN(atomic_addr)(&g_sites[3 * i + 1], sg.y);
    FN(atomic_addr)(&g_sites[3 * i + 2], sg.z);
    dLdx = FN(sub)(dLdx, sg);
  }
  R dLdtq = FN(dot)(dLdx, d);
  R dtn = (f->dp >= (R)0 || t_near0 >= f->t_surf) ? dLdtq : (R)0;
  R dts = (f->dp < (R)0 && t_near0 < f->t_surf) ? dLdtq : (R)0;
  dts += dLdt_surf_in;
  if (fabs((double)dts) > 0.0) {
    V3 a, b;
    R c;
    FN(ray_plane_bwd)(o, d, p, n, f->height, eps, &a, &b, &c);
    *dLdp = FN(add)(*dLdp, FN(scl)(a, dts));
    *dLdn = FN(add)(*dLdn, FN(scl)(b, dts));
    dLdheight += c * dts;
  }
  /* height stage */
  R tq0 = f->dp0 >= (R)0 ? t_near0 : FN(rmax)(t_near0, f->t_surf0);
  V3 x0 = FN(add)(o, FN(scl)(d, tq0));
  V3 dLdx0 = FN(mk)(0, 0, 0);
  R gdp0_in = dLdheight * f->height;
  for (int i = 0; i < T; i++) {
    V3 s = FN(ld3)(sites + 3 * i);
    V3 df = FN(sub)(x0, s);
    R w = FN(rexp)(-(R)10 * (FN(dot)(df, df) / r2));
    /* the reference tests against the colour-stage weight sum here */
    if (w / f->wsum < cut) continue;
    V3 dwds = FN(dvs)(FN(scl)(df, (R)2 * (R)10 * w), r2);
    FN(atomic_addr)(&g_height[i], w * dLdheight / f->wsum0);
    R gdp = dLdheight * (R)rgbh[4 * i + 3];
    V3 sg = FN(dvs)(FN(scl)(dwds, gdp - gdp0_in), f->wsum0);
    FN(atomic_addr)(&g_sites[3 * i + 0], sg.x);
    FN(atomic_addr)(&g_sites[3 * i + 1], sg.y);
    FN(atomic_addr)(&g_sites[3 * i + 2], sg.z);
    dLdx0 = FN(sub)(dLdx0, sg);
  }
  R dLdtq0 = FN(dot)(dLdx0, d);
  dtn += (f->dp >= (R)0 || t_near0 >= f->t_surf0) ? dLdtq0 : (R)0;
  R dts0 = (f->dp < (R)0 && t_near0 < f->t_surf0) ? dLdtq0 : (R)0;
  if (fabs((double)dts0) > 0.0) {
    V3 a, b;
    R c;
    FN(ray_plane_bwd)(o, d, p, n, (R)0, eps, &a, &b, &c);
    *dLdp = FN(add)(*dLdp, FN(scl)(a, dts0));
    *dLdn = FN(add)(*dLdn, FN(scl)(b, dts0));
  }
  *dLdtn = dtn;
}

/* The ray of a lane (rasterize.py:919-955; camera.py:19-24). */
static inline void FN(lane_ray)(const pf_camera* cam, int tile, int lane, int* pi, int* pj,
                                int* oob, V3* o, V3* d) {
  int tiles_h = 1 + (cam->height - 1) / PF_TILE_WIDTH;
  int tile_i = tile % tiles_h, tile_j = tile / tiles_h;
  int i = tile_i * PF_TILE_WIDTH + lane % PF_TILE_WIDTH;
  int j = tile_j * PF_TILE_WIDTH + lane / PF_TILE_WIDTH;
  *oob = (i >= cam->height || j >= cam->width);
  if (i > cam->height - 1) i = cam->height - 1;
  if (j > cam->width - 1) j = cam->width - 1;
  *pi = i;
  *pj = j;
  if (cam->ray_maps == NULL) {
    V3 right = FN(ld3)(cam->right), up = FN(ld3)(cam->up);
    V3 fwd = FN(nrm)(FN(crs)(up, right));
    R x = (R)2 * (R)j / ((R)cam->width - (R)1) - (R)1;
    R y = (R)1 - (R)2 * (R)i / ((R)cam->height - (R)1);
    V3 rd = FN(add)(FN(add)(FN(scl)(right, x), FN(scl)(up, y)), fwd);
    *d = FN(dvs)(rd, FN(len)(rd));
    *o = FN(ld3)(cam->eye);
  } else {
    const float* rm = cam->ray_maps + ((size_t)i * cam->width + j) * 6;
    *o = FN(ld3)(rm);
    V3 rd = FN(ld3)(rm + 3);
    *d = FN(dvs)(rd, FN(len)(rd));
  }
}

/* Outputs of the FP64/FP32 oracle forward are of type R. */
typedef struct {
  int32_t* tile_early_stop;
  R* color;
  R* log_t;
  R* normal_distance;
  R* normal;
  R* quantile_depths;
  R* err;
  R* contrib;
  R* point_err;
} FN(fwd_out);

/* Face constraint of a site against its e-th neighbour as the kernels see
 * it: neighbour centre c + fp16 delta, neighbour radius fp16 (training
 * encoding, rasterize.py:1059-1064). With no encoded deltas the neighbour's
 * full-precision sphere is used (exact geometry, for derivative checks). */
static inline void FN(face_geom)(const pf_scene* sc, int64_t e, V3 c, V3* a, R* ra) {
  if (sc->adjacency_diff == NULL) {
    const float* s = sc->spheres + 4 * (size_t)sc->adjacency[e];
    *a = FN(ld3)(s);
    *ra = (R)s[3];
    return;
  }
  const uint16_t* h = sc->adjacency_diff + 4 * e;
  *a = FN(add)(c, FN(mk)((R)orc_f16_to_f32(h[0]), (R)orc_f16_to_f32(h[1]),
                          (R)orc_f16_to_f32(h[2])));
  *ra = (R)orc_f16_to_f32(h[3]);
}

void FN(orc_raster_fwd)(const pf_raster_fwd_args* A, FN(fwd_out) * out) {
  const pf_camera* cam = &A->cam;
  const pf_scene* sc = &A->scene;
  const int H = cam->height, W = cam->width;
  const int n_tiles = A->tiles.n_tiles;
  const int T = sc->n_texels;
  const int Q = A->depth_quantiles ? A->n_quantiles : 0;
  const R HWf = (R)(float)(H * W);
  const R thr = (R)A->transmittance_threshold;

#pragma omp parallel for schedule(dynamic, 4)
  for (int tile = 0; tile < n_tiles; tile++) {
    int pi[64], pj[64], oob[64];
    V3 ro[64], rd[64];
    R log_t[64], nd[64], err[64];
    V3 rgb[64], nacc[64], gt[64];
    int qidx[64];
    for (int l = 0; l < 64; l++) {
      FN(lane_ray)(cam, tile, l, &pi[l], &pj[l], &oob[l], &ro[l], &rd[l]);
      log_t[l] = 0; nd[l] = 0; err[l] = 0;
      rgb[l] = FN(mk)(0, 0, 0); nacc[l] = FN(mk)(0, 0, 0);
      qidx[l] = 0;
      gt[l] = A->ray_gt ? FN(ld3)(A->ray_gt + 3 * ((size_t)pi[l] * W + pj[l])) : FN(mk)(0, 0, 0);
    }
    const int64_t start = A->tiles.offsets[tile];
    const int total = (int)(A->tiles.offsets[tile + 1] - start);
    int early = total;
    for (int k = 0; k < total; k++) {
      R max_t = FN(rexp)(log_t[0]);
      for (int l = 1; l < 64; l++) max_t = FN(rmax)(max_t, FN(rexp)(log_t[l]));
      if (max_t < thr) {
        early = k;
        break;
      }
      const int site = A->tiles.sites[start + k];
      const float* sp = sc->spheres + 4 * (size_t)site;
      const float* ns = sc->nsigmas + 4 * (size_t)site;
      V3 c = FN(ld3)(sp);
      R r = (R)sp[3];
      V3 pn = FN(ld3)(ns);
      R sigma = (R)ns[3];
      FN(sphere_hit) hs[64];
      int any = 0;
      for (int l = 0; l < 64; l++) {
        hs[l] = FN(ray_sphere)(ro[l], rd[l], c, r);
        any |= hs[l].hit;
      }
      if (!any) continue;
      const int a0 = sc->adjacency_offsets[site], a1 = sc->adjacency_offsets[site + 1];
      const float* tsites = sc->texel_sites + (size_t)site * T * 3;
      const float* trgbh = sc->texel_rgbh + (size_t)site * T * 4;
      for (int l = 0; l < 64; l++) {
        R tn = hs[l].t_near, tf = hs[l].t_far;
        if (hs[l].hit) {
          for (int e = a0; e < a1; e++) {
            V3 a; R ra, t, dp;
            FN(face_geom)(sc, e, c, &a, &ra);
            FN(ray_pface)(ro[l], rd[l], c, r, a, ra, &t, &dp);
            tf = dp >= (R)0 ? FN(rmin)(t, tf) : tf;
            tn = dp < (R)0 ? FN(rmax)(t, tn) : tn;
          }
        }
        FN(plane_fwd_out) pf = FN(plane_fwd)(ro[l], rd[l], tn, c, pn, r, !oob[l] && hs[l].hit,
                                             tsites, trgbh, T);
        tf = pf.dp >= (R)0 ? FN(rmin)(pf.t_surf, tf) : tf;
        tn = pf.dp < (R)0 ? FN(rmax)(pf.t_surf, tn) : tn;
        R dt = tf - tn;
        if (hs[l].hit && dt > (R)0) {
          R dlt = -sigma * dt;
          R alpha = (R)1 - FN(rexp)(dlt);
          R trans = FN(rexp)(log_t[l]);
          if (A->loss_accumulation) {
            V3 df = FN(sub)(pf.rgb, gt[l]);
            err[l] += FN(dot)(df, df) * alpha * trans;
          }
          rgb[l] = FN(add)(rgb[l], FN(scl)(pf.rgb, alpha * trans));
          log_t[l] += dlt;
          FN(atomic_addr)(&out->contrib[site], alpha * trans / HWf);
          if (A->return_point_err) {
            V3 df = FN(sub)(pf.rgb, gt[l]);
            R l1 = fabs((double)df.x) + fabs((double)df.y) + fabs((double)df.z);
            FN(atomic_addr)(&out->point_err[site], alpha * trans * l1 / HWf);
          }
          R ndv = FN(dot)(pn, rd[l]);
          if (ndv > (R)0) nd[l] += (ndv * ndv) * alpha * trans;
          nacc[l] = FN(add)(nacc[l], FN(scl)(pn, alpha * trans));
          R next_t = FN(rexp)(log_t[l]);
          if (Q > 0) {
            const float* qs = A->depth_quantiles + ((size_t)pi[l] * W + pj[l]) * Q;
            while (qidx[l] < Q && next_t < (R)qs[qidx[l]]) {
              R qd = tn + FN(rlog)(trans / (R)qs[qidx[l]]) / sigma;
              if (!oob[l]) out->quantile_depths[((size_t)pi[l] * W + pj[l]) * Q + qidx[l]] = qd;
              qidx[l]++;
            }
          }
        }
      }
    }
    out->tile_early_stop[tile] = early;
    for (int l = 0; l < 64; l++) {
      if (oob[l]) continue;
      size_t px = (size_t)pi[l] * W + pj[l];
      R tr = FN(rexp)(log_t[l]);
      (void)tr; /* background is black */
      out->color[3 * px + 0] = rgb[l].x;
      out->color[3 * px + 1] = rgb[l].y;
      out->color[3 * px + 2] = rgb[l].z;
      out->log_t[px] = log_t[l];
      out->normal_distance[px] = nd[l];
      out->normal[3 * px + 0] = nacc[l].x;
      out->normal[3 * px + 1] = nacc[l].y;
      out->normal[3 * px + 2] = nacc[l].z;
      if (A->loss_accumulation && out->err) {
        out->err[px] = err[l] + FN(dot)(gt[l], gt[l]) * tr;
      }
    }
  }
}

/* Gradient outputs of type R. */
typedef struct {
  R* grad_spheres;
  R* grad_nsigmas;
  R* grad_texel_sites;
  R* grad_texel_rgb;
  R* grad_texel_height;
  /* optional [N]: 1 where a ray meets the site ill-conditioned -- tangent to
   * its sphere (|r^2 - |oc x d|^2| < 1e-6 r^2) or with a clipped interval
   * whose length is within 1e-5 (1 + t_far) of zero. There the gradient
   * grows like 1/sqrt of that margin, so one-ulp differences between
   * implementations flip whole contributions; the site and its neighbours
   * are flagged. */
  uint8_t* fragile;
} FN(bwd_out);

void FN(orc_raster_bwd)(const pf_raster_bwd_args* A, FN(bwd_out) * out) {
  const pf_camera* cam = &A->cam;
  const pf_scene* sc = &A->scene;
  const int H = cam->height, W = cam->width;
  const int n_tiles = A->tiles.n_tiles;
  const int T = sc->n_texels;
  const int Q = A->depth_quantiles ? A->n_quantiles : 0;
  const R HWf = (R)(float)(H * W);
  const int exact = A->exact_gradients;
  const R qeps = exact ? (R)0 : (R)1e-6;

#pragma omp parallel for schedule(dynamic, 4)
  for (int tile = 0; tile < n_tiles; tile++) {
    int pi[64], pj[64], oob[64];
    V3 ro[64], rd[64];
    R log_t[64], dLdlog_t[64], gnd[64];
    V3 grgb[64], gn[64];
    int qidx[64];
    for (int l = 0; l < 64; l++) {
      FN(lane_ray)(cam, tile, l, &pi[l], &pj[l], &oob[l], &ro[l], &rd[l]);
      size_t px = (size_t)pi[l] * W + pj[l];
      log_t[l] = (R)A->log_t[px];
      qidx[l] = Q - 1;
      grgb[l] = FN(ld3)(A->grad_color + 3 * px);
      gn[l] = FN(ld3)(A->grad_normal + 3 * px);
      gnd[l] = (R)A->grad_normal_distance[px];
      /* background is black: dL/dT_final from colour is 0 */
      dLdlog_t[l] = (R)A->grad_log_t[px];
    }
    const int64_t start = A->tiles.offsets[tile];
    const int early = A->tile_early_stop[tile];
    for (int k = early - 1; k >= 0; k--) {
      const int site = A->tiles.sites[start + k];
      const float* sp = sc->spheres + 4 * (size_t)site;
      const float* ns = sc->nsigmas + 4 * (size_t)site;
      V3 c = FN(ld3)(sp);
      R r = (R)sp[3];
      V3 pn = FN(ld3)(ns);
      R sigma = (R)ns[3];
      FN(sphere_hit) hs[64];
      int any = 0;
      for (int l = 0; l < 64; l++) {
        hs[l] = FN(ray_sphere)(ro[l], rd[l], c, r);
        any |= hs[l].hit;
      }
      if (!any) continue;
      const int a0 = sc->adjacency_offsets[site], a1 = sc->adjacency_offsets[site + 1];
      const int n_adj = a1 - a0;
      const float* tsites = sc->texel_sites + (size_t)site * T * 3;
      const float* trgbh = sc->texel_rgbh + (size_t)site * T * 4;
      R* g_ts = out->grad_texel_sites + (size_t)site * T * 3;
      R* g_tr = out->grad_texel_rgb + (size_t)site * T * 3;
      R* g_th = out->grad_texel_height + (size_t)site * T;
      for (int l = 0; l < 64; l++) {
        R tn = hs[l].t_near, tf = hs[l].t_far;
        int tn_id = tn < (R)1e-8 ? -2 : -1;
        int tf_id = tf < (R)1e-8 ? -2 : -1;
        if (hs[l].hit) {
          for (int e = 0; e < n_adj; e++) {
            V3 a; R ra, t, dp;
            FN(face_geom)(sc, a0 + e, c, &a, &ra);
            FN(ray_pface)(ro[l], rd[l], c, r, a, ra, &t, &dp);
            tf_id = (dp >= (R)0 && t < tf) ? e : tf_id;
            tf = dp >= (R)0 ? FN(rmin)(t, tf) : tf;
            tn_id = (dp < (R)0 && t > tn) ? e : tn_id;
            tn = dp < (R)0 ? FN(rmax)(t, tn) : tn;
          }
        }
        const R tn0 = tn;
        const int tn0_id = tn_id, tf0_id = tf_id;
        FN(plane_fwd_out) pf = FN(plane_fwd)(ro[l], rd[l], tn, c, pn, r, !oob[l] && hs[l].hit,
                                             tsites, trgbh, T);
        tf_id = (pf.dp >= (R)0 && pf.t_surf < tf) ? n_adj : tf_id;
        tf = pf.dp >= (R)0 ? FN(rmin)(pf.t_surf, tf) : tf;
        tn_id = (pf.dp < (R)0 && pf.t_surf > tn) ? n_adj : tn_id;
        tn = pf.dp < (R)0 ? FN(rmax)(pf.t_surf, tn) : tn;
        R dt = tf - tn;
        if (out->fragile && !oob[l]) {
          V3 oc = FN(sub)(ro[l], c);
          V3 cr = FN(crs)(oc, rd[l]);
          R h2 = r * r - FN(dot)(cr, cr);
          int frag = fabs((double)h2) < g_frag_h2_rel * (double)(r * r) +
                                            g_frag_h2_abs * (double)r * (double)FN(len)(oc);
          if (hs[l].hit && fabs((double)dt) < g_frag_dt_rel * (1.0 + fabs((double)tf))) frag = 1;
          if (frag) {
            out->fragile[site] = 1;
            for (int e = a0; e < a1; e++) out->fragile[sc->adjacency[e]] = 1;
          }
        }
        R dLdtn = 0, dLdtf = 0, dLdr = 0, dLdsig = 0;
        V3 dLdc = FN(mk)(0, 0, 0), dLdnorm = FN(mk)(0, 0, 0), dLdcol = FN(mk)(0, 0, 0);
        const int active = !oob[l] && hs[l].hit && dt > (R)0;
        if (active) {
          size_t px = (size_t)pi[l] * W + pj[l];
          R dlt = -sigma * dt;
          log_t[l] -= dlt;
          R alpha = (R)1 - FN(rexp)(dlt);
          R trans = FN(rexp)(log_t[l]);
          R ndv = FN(dot)(pn, rd[l]);
          R dLdlog_t_i = dLdlog_t[l];
          R dLddlt = dLdlog_t[l];
          R gc = (R)A->grad_contrib[site];
          R cdot = FN(dot)(grgb[l], pf.rgb);
          R ndt = exact ? ndv * ndv : ndv;
          R dLdalpha = cdot * trans;
          if (ndv > (R)0) dLdalpha += gnd[l] * ndt * trans;
          dLdalpha += gc * trans / HWf;
          dLdalpha += FN(dot)(gn[l], pn) * trans;
          R dLdtrans = cdot * alpha;
          if (ndv > (R)0) dLdtrans += gnd[l] * ndt * alpha;
          dLdtrans += gc * alpha / HWf;
          dLdtrans += FN(dot)(gn[l], pn) * alpha;
          dLdcol = FN(scl)(grgb[l], trans * alpha);
          dLdlog_t_i += dLdtrans * FN(rexp)(log_t[l]);
          dLddlt += -dLdalpha * FN(rexp)(dlt);
          dLdsig += dLddlt * -dt;
          dLdtn += dLddlt * sigma;
          dLdtf += dLddlt * -sigma;
          if (Q > 0) {
            const float* qs = A->depth_quantiles + px * Q;
            /* no incoming quantile gradient (NULL): the crossings still
             * advance, each adding zero */
            const float* gq = A->grad_quantile_depths ? A->grad_quantile_depths + px * Q : NULL;
            while (qidx[l] >= 0 && trans > (R)qs[qidx[l]]) {
              R g = gq ? (R)gq[qidx[l]] : (R)0;
              dLdsig += -g * FN(rlog)(trans / (R)qs[qidx[l]]) / (sigma * sigma + qeps);
              dLdtn += g;
              dLdlog_t_i += g / (sigma + qeps);
              qidx[l]--;
            }
          }
          dLdnorm = FN(add)(dLdnorm, FN(scl)(gn[l], alpha * trans));
          if (ndv > (R)0) {
            R s = exact ? (R)2 * ndv : (R)1;
            dLdnorm = FN(add)(dLdnorm, FN(scl)(rd[l], gnd[l] * alpha * trans * s));
          }
          dLdlog_t[l] = dLdlog_t_i;
        }
        R dLdts;
        if (tn_id == n_adj) dLdts = dLdtn;
        else if (tf_id == n_adj) dLdts = dLdtf;
        else dLdts = 0;
        V3 pdc, pdn;
        R dLdtn_q;
        FN(plane_bwd)(ro[l], rd[l], tn0, c, pn, r, active, tsites, trgbh, &pf, dLdcol, dLdts, T,
                      exact, g_ts, g_tr, g_th, &pdc, &pdn, &dLdtn_q);
        dLdc = FN(add)(dLdc, pdc);
        dLdnorm = FN(add)(dLdnorm, pdn);
        if (tn_id != n_adj) dLdtn_q += dLdtn;
        if (active) {
          for (int side = 0; side < 2; side++) {
            int tid = side == 0 ? tn0_id : tf0_id;
            R dLdt = side == 0 ? dLdtn_q : dLdtf;
            /* the exact mode sends the far end to the plane only when the
             * plane is the far constraint */
            if (exact && side == 1 && tf_id == n_adj) continue;
            if (tid == -1) {
              FN(sphere_grad) sg = FN(ray_sphere_bwd)(ro[l], rd[l], c, r);
              if (side == 0) {
                dLdc = FN(add)(dLdc, FN(scl)(sg.dtn_dc, dLdt));
                dLdr += dLdt * sg.dtn_dr;
              } else {
                dLdc = FN(add)(dLdc, FN(scl)(sg.dtf_dc, dLdt));
                dLdr += dLdt * sg.dtf_dr;
              }
            } else if (tid >= 0 && tid < n_adj) {
              int adj = sc->adjacency[a0 + tid];
              const float* asp = sc->spheres + 4 * (size_t)adj;
              V3 ac = FN(ld3)(asp);
              R ar = (R)asp[3];
              V3 dtdc, dtda;
              R dtdr, dtdra;
              FN(ray_pface_bwd)(ro[l], rd[l], c, r, ac, ar, &dtdc, &dtdr, &dtda, &dtdra);
              dLdc = FN(add)(dLdc, FN(scl)(dtdc, dLdt));
              dLdr += dLdt * dtdr;
              R* ga = out->grad_spheres + 4 * (size_t)adj;
              FN(atomic_addr)(&ga[0], dLdt * dtda.x);
              FN(atomic_addr)(&ga[1], dLdt * dtda.y);
              FN(atomic_addr)(&ga[2], dLdt * dtda.z);
              FN(atomic_addr)(&ga[3], dLdt * dtdra);
            }
          }
          R* gs = out->grad_spheres + 4 * (size_t)site;
          FN(atomic_addr)(&gs[0], dLdc.x);
          FN(atomic_addr)(&gs[1], dLdc.y);
          FN(atomic_addr)(&gs[2], dLdc.z);
          FN(atomic_addr)(&gs[3], dLdr);
          R* gnsig = out->grad_nsigmas + 4 * (size_t)site;
          FN(atomic_addr)(&gnsig[0], dLdnorm.x);
          FN(atomic_addr)(&gnsig[1], dLdnorm.y);
          FN(atomic_addr)(&gnsig[2], dLdnorm.z);
          FN(atomic_addr)(&gnsig[3], dLdsig);
        }
      }
    }
  }
}


/* Benchmark forward (benchmark_kernel, rasterize.py:614-734): colour only,
 * per-lane stop at T < threshold, faces from the (c_adj - c, pm_adj - pm)
 * binary16 encoding of benchmark.py:297-302 via ray_pface_intersect_diff
 * (rendering_math.py:85-98), a lane leaves the face loop once its interval
 * is empty, plane_intersection_fwd_local (rasterize.py:67-138). */
void FN(orc_raster_rgb)(const pf_camera* cam, const pf_scene* sc, const int64_t* offsets, const int32_t* sites,
                        int n_tiles, float threshold, R* color) {
  const int W = cam->width;
  const int T = sc->n_texels;
#pragma omp parallel for schedule(dynamic, 4)
  for (int tile = 0; tile < n_tiles; tile++) {
    for (int l = 0; l < 64; l++) {
      int pi, pj, oob;
      V3 o, d;
      FN(lane_ray)(cam, tile, l, &pi, &pj, &oob, &o, &d);
      if (oob) continue;
      V3 rgb = FN(mk)(0, 0, 0);
      R log_t = 0;
      for (int64_t k = offsets[tile]; k < offsets[tile + 1]; k++) {
        R trans = FN(rexp)(log_t);
        if (trans < (R)threshold) break;
        const int site = sites[k];
        const float* sp = sc->spheres + 4 * (size_t)site;
        const float* ns = sc->nsigmas + 4 * (size_t)site;
        V3 c = FN(ld3)(sp);
        R r = (R)sp[3];
        FN(sphere_hit) h = FN(ray_sphere)(o, d, c, r);
        if (!h.hit) continue;
        R tn = h.t_near, tf = h.t_far;
        for (int e = sc->adjacency_offsets[site]; e < sc->adjacency_offsets[site + 1]; e++) {
          const uint16_t* q = sc->adjacency_diff + 4 * (size_t)e;
          V3 df = FN(mk)((R)orc_f16_to_f32(q[0]), (R)orc_f16_to_f32(q[1]), (R)orc_f16_to_f32(q[2]));
          R pm = (R)orc_f16_to_f32(q[3]);
          R dp = FN(dot)(d, df);
          R t = (pm - FN(dot)(o, df)) / dp;
          tf = dp >= (R)0 ? FN(rmin)(t, tf) : tf;
          tn = dp < (R)0 ? FN(rmax)(t, tn) : tn;
          if (tn > tf) break;
        }
        if (tn > tf) continue;
        FN(plane_fwd_out) pf = FN(plane_fwd)(o, d, tn, c, FN(ld3)(ns), r, 1,
                                             sc->texel_sites + (size_t)site * T * 3,
                                             sc->texel_rgbh + (size_t)site * T * 4, T);
        tf = pf.dp >= (R)0 ? FN(rmin)(pf.t_surf, tf) : tf;
        tn = pf.dp < (R)0 ? FN(rmax)(pf.t_surf, tn) : tn;
        R dt = tf - tn;
        if (dt > (R)0) {
          R dlt = -(R)ns[3] * dt;
          R alpha = (R)1 - FN(rexp)(dlt);
          rgb = FN(add)(rgb, FN(scl)(pf.rgb, alpha * trans));
          log_t += dlt;
        }
      }
      size_t px = (size_t)pi * W + pj;
      color[3 * px + 0] = rgb.x;
      color[3 * px + 1] = rgb.y;
      color[3 * px + 2] = rgb.z;
    }
  }
}


/* Visualisation (visualization_kernel, rasterize.py:739-890): colour over a
 * background, the first depth where T falls below depth_q (0 if never),
 * alpha-weighted normals, alpha and the number of list entries each ray
 * looked at. Per ray: stop after max_inter entries or at T < threshold
 * (the entry that stops it is counted); faces from the training encoding
 * (c_adj - c, r_adj) (prefetch_adjacency_kernel), leaving the face loop once
 * the interval is empty. */
typedef struct {
  R* color;
  R* depth;
  R* normal;
  R* alpha;
  int32_t* inter;
} FN(vis_out);

void FN(orc_raster_vis)(const pf_camera* cam, const pf_scene* sc, const int64_t* offsets, const int32_t* sites,
                        int n_tiles, float threshold, int max_inter, float depth_q, const float* bkgd,
                        FN(vis_out) * out) {
  const int W = cam->width;
  const int T = sc->n_texels;
#pragma omp parallel for schedule(dynamic, 4)
  for (int tile = 0; tile < n_tiles; tile++) {
    for (int l = 0; l < 64; l++) {
      int pi, pj, oob;
      V3 o, d;
      FN(lane_ray)(cam, tile, l, &pi, &pj, &oob, &o, &d);
      if (oob) continue;
      V3 rgb = FN(mk)(0, 0, 0), nacc = FN(mk)(0, 0, 0);
      R log_t = 0, depth = 0;
      int found = 0, nhit = 0;
      const int total = (int)(offsets[tile + 1] - offsets[tile]);
      for (int k = 0; k < total; k++) {
        nhit++;
        if (k >= max_inter) break;
        R trans = FN(rexp)(log_t);
        if (trans < (R)threshold) break;
        const int site = sites[offsets[tile] + k];
        const float* sp = sc->spheres + 4 * (size_t)site;
        const float* ns = sc->nsigmas + 4 * (size_t)site;
        V3 c = FN(ld3)(sp);
        R r = (R)sp[3];
        FN(sphere_hit) h = FN(ray_sphere)(o, d, c, r);
        if (!h.hit) continue;
        R tn = h.t_near, tf = h.t_far;
        for (int e = sc->adjacency_offsets[site]; e < sc->adjacency_offsets[site + 1]; e++) {
          V3 a;
          R ra, t, dp;
          FN(face_geom)(sc, e, c, &a, &ra);
          FN(ray_pface)(o, d, c, r, a, ra, &t, &dp);
          tf = dp >= (R)0 ? FN(rmin)(t, tf) : tf;
          tn = dp < (R)0 ? FN(rmax)(t, tn) : tn;
          if (tn > tf) break;
        }
        if (tn > tf) continue;
        V3 pn = FN(ld3)(ns);
        R sigma = (R)ns[3];
        FN(plane_fwd_out) pf = FN(plane_fwd)(o, d, tn, c, pn, r, 1, sc->texel_sites + (size_t)site * T * 3,
                                             sc->texel_rgbh + (size_t)site * T * 4, T);
        tf = pf.dp >= (R)0 ? FN(rmin)(pf.t_surf, tf) : tf;
        tn = pf.dp < (R)0 ? FN(rmax)(pf.t_surf, tn) : tn;
        R dt = tf - tn;
        if (dt > (R)0) {
          R dlt = -sigma * dt;
          R alpha = (R)1 - FN(rexp)(dlt);
          rgb = FN(add)(rgb, FN(scl)(pf.rgb, alpha * trans));
          nacc = FN(add)(nacc, FN(scl)(pn, alpha * trans));
          log_t += dlt;
          if (!found && FN(rexp)(log_t) < (R)depth_q) {
            depth = tn + FN(rlog)(trans / (R)depth_q) / sigma;
            found = 1;
          }
        }
      }
      const size_t px = (size_t)pi * W + pj;
      const R tr = FN(rexp)(log_t);
      out->color[3 * px + 0] = rgb.x + (R)bkgd[0] * tr;
      out->color[3 * px + 1] = rgb.y + (R)bkgd[1] * tr;
      out->color[3 * px + 2] = rgb.z + (R)bkgd[2] * tr;
      out->depth[px] = depth;
      out->normal[3 * px + 0] = nacc.x;
      out->normal[3 * px + 1] = nacc.y;
      out->normal[3 * px + 2] = nacc.z;
      out->alpha[px] = (R)1 - tr;
      out->inter[px] = nhit;
    }
  }
}


/* ------------------------------------------------------------ ray tracer
 * RayTracer.benchmark_kernel (raytrace.py:115-241): per pixel, walk the power
 * cells from a start cell; the next cell is the neighbour whose face the ray
 * leaves through first (argmin t over faces with d.n >= 0, first in list order
 * on ties); a cell shades when the ray meets its ball (not within 4 r of the
 * centre), its interval is non-empty and sigma >= 1e-3. Faces come from the
 * (c_adj - c, pm_adj - pm) binary16 encoding, or from the spheres when no
 * encoding is given (exact geometry). Added outputs: log T, the alpha-weighted
 * normal, the depth where T first drops below depth_q (-1 if never), the
 * visited cells (first K) and their count. */
typedef struct {
  R* color;
  R* log_t;
  R* normal;
  R* depth;
  int32_t* q_step;  /* walk step of the depth crossing, -1 if none */
  R* q_sigma;
  int32_t* cells;   /* [P, K] or NULL */
  int32_t* n_cells; /* [P] */
  uint8_t* fragile; /* [P] or NULL: the ray crosses an ill-conditioned cell */
  int K;
} FN(trace_out);

static inline void FN(trace_face)(const pf_scene* sc, int64_t e, V3 c, R r, V3* n, R* pm) {
  if (sc->adjacency_diff) {
    const uint16_t* q = sc->adjacency_diff + 4 * e;
    *n = FN(mk)((R)orc_f16_to_f32(q[0]), (R)orc_f16_to_f32(q[1]), (R)orc_f16_to_f32(q[2]));
    *pm = (R)orc_f16_to_f32(q[3]);
  } else {
    const float* a = sc->spheres + 4 * (size_t)sc->adjacency[e];
    V3 ac = FN(ld3)(a);
    R ra = (R)a[3];
    *n = FN(sub)(ac, c);
    *pm = (R)0.5 * (FN(dot)(ac, ac) - ra * ra) - (R)0.5 * (FN(dot)(c, c) - r * r);
  }
}

/* The ray of pixel (i, j) (camera.py:19-24 / ray_maps). */
static inline void FN(pixel_ray)(const pf_camera* cam, int i, int j, V3* o, V3* d) {
  if (cam->ray_maps == NULL) {
    V3 right = FN(ld3)(cam->right), up = FN(ld3)(cam->up);
    V3 fwd = FN(nrm)(FN(crs)(up, right));
    R x = (R)2 * (R)j / ((R)cam->width - (R)1) - (R)1;
    R y = (R)1 - (R)2 * (R)i / ((R)cam->height - (R)1);
    V3 rd = FN(add)(FN(add)(FN(scl)(right, x), FN(scl)(up, y)), fwd);
    *d = FN(dvs)(rd, FN(len)(rd));
    *o = FN(ld3)(cam->eye);
  } else {
    const float* rm = cam->ray_maps + ((size_t)i * cam->width + j) * 6;
    *o = FN(ld3)(rm);
    V3 rd = FN(ld3)(rm + 3);
    *d = FN(dvs)(rd, FN(len)(rd));
  }
}

/* A cell of the walk is ill-conditioned when a float32 rounding of its inputs
 * can change which bound binds, whether the ball is met, or which cell comes
 * next, or is amplified past 1e-4: a grazing ball (h^2 = r^2 - |oc x d|^2 is
 * computed with an absolute error near 1e-7 |oc| r, and the interval ends move
 * like 1/h, so h^2 < 1e-3 |oc| r counts), a centre within a hair of the 4 r
 * exclusion, a face nearly parallel to the ray, or two candidate bounds
 * (faces and the ball) within a hair of each other on either end of the
 * interval. Rays through such cells are left out of gradient comparisons
 * between precisions. */
static inline int FN(trace_fragile)(const pf_scene* sc, V3 o, V3 d, V3 c, R r, FN(sphere_hit) h, int a0, int a1,
                                    R tn, R tf, R pt_far) {
  const V3 cr = FN(crs)(FN(sub)(o, c), d);
  const double h2 = (double)r * (double)r - (double)FN(dot)(cr, cr);
  const double dist = (double)FN(len)(FN(sub)(c, o));
  const double tol = 1e-5 * (1.0 + fabs((double)tf));
  if (fabs(h2) < 1e-3 * dist * (double)r || fabs(dist - 4.0 * (double)r) < 1e-5 * (double)r) return 1;
  if (h.hit && fabs((double)(tf - tn)) < tol) return 1;
  int near_ties = h.hit && fabs((double)(h.t_near - tn)) < tol;
  int far_ties = h.hit && fabs((double)(h.t_far - tf)) < tol;
  int next_ties = 0;
  for (int e = a0; e < a1; e++) {
    V3 n;
    R pm;
    FN(trace_face)(sc, e, c, r, &n, &pm);
    const R dp = FN(dot)(d, n);
    if (fabs((double)dp) < 1e-6 * (double)FN(len)(n)) return 1;
    const R t = (pm - FN(dot)(o, n)) / dp;
    if (dp >= (R)0) {
      far_ties += fabs((double)(t - tf)) < tol;
      next_ties += fabs((double)(t - pt_far)) < tol;
    } else {
      near_ties += fabs((double)(t - tn)) < tol;
    }
  }
  return near_ties > 1 || far_ties > 1 || next_ties > 1;
}

void FN(orc_trace_fwd)(const pf_camera* cam, const pf_scene* sc, const int32_t* start_sites, int start_site,
                       float threshold, float depth_q, int max_steps, FN(trace_out) * out) {
  const int H = cam->height, W = cam->width, T = sc->n_texels;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t px = 0; px < (int64_t)H * W; px++) {
    const int i = (int)(px / W), j = (int)(px % W);
    V3 o, d;
    FN(pixel_ray)(cam, i, j, &o, &d);
    int prim = start_sites ? start_sites[px] : start_site;
    R log_t = 0, depth = -1, qsig = 0;
    int qstep = -1, step = 0, frag = 0;
    V3 rgb = FN(mk)(0, 0, 0), nacc = FN(mk)(0, 0, 0);
    while (1) {
      R trans = FN(rexp)(log_t);
      if (trans < (R)threshold || prim == 0x7FFFFFFF || prim < 0 || step >= max_steps) break;
      if (out->cells && step < out->K) out->cells[px * out->K + step] = prim;
      const float* sp = sc->spheres + 4 * (size_t)prim;
      const float* ns = sc->nsigmas + 4 * (size_t)prim;
      V3 c = FN(ld3)(sp);
      R r = (R)sp[3];
      FN(sphere_hit) h = FN(ray_sphere)(o, d, c, r);
      int hit = h.hit;
      if (FN(len)(FN(sub)(c, o)) < (R)4 * r) hit = 0;
      R tn = h.t_near, tf = h.t_far;
      int next = 0x7FFFFFFF;
      R pt_far = (R)1e10;
      for (int e = sc->adjacency_offsets[prim]; e < sc->adjacency_offsets[prim + 1]; e++) {
        V3 n;
        R pm;
        FN(trace_face)(sc, e, c, r, &n, &pm);
        R dp = FN(dot)(d, n);
        R t = (pm - FN(dot)(o, n)) / dp;
        if (dp >= (R)0 && t < pt_far) {
          next = sc->adjacency[e];
          pt_far = t;
        }
        tf = dp >= (R)0 ? FN(rmin)(t, tf) : tf;
        tn = dp < (R)0 ? FN(rmax)(t, tn) : tn;
      }
      R sigma = (R)ns[3];
      if (out->fragile && FN(trace_fragile)(sc, o, d, c, r, h, sc->adjacency_offsets[prim],
                                            sc->adjacency_offsets[prim + 1], tn, tf, pt_far))
        frag = 1;
      if (!hit || tn > tf || sigma < (R)1e-3) {
        prim = next;
        step++;
        continue;
      }
      FN(plane_fwd_out) pf = FN(plane_fwd)(o, d, tn, c, FN(ld3)(ns), r, 1, sc->texel_sites + (size_t)prim * T * 3,
                                           sc->texel_rgbh + (size_t)prim * T * 4, T);
      const double tsel = pf.dp >= (R)0 ? (double)tf : (double)tn;
      tf = pf.dp >= (R)0 ? FN(rmin)(pf.t_surf, tf) : tf;
      tn = pf.dp < (R)0 ? FN(rmax)(pf.t_surf, tn) : tn;
      R dt = tf - tn;
      if (out->fragile) {
        /* surface nearly parallel to the ray, its bound nearly tied with the
         * other bound on that end, a near-empty interval, or T after this
         * cell within a hair of the depth quantile or the early-stop level */
        const double tol = 1e-5 * (1.0 + fabs((double)tf));
        frag |= fabs((double)pf.dp) < 1e-6 || fabs((double)pf.t_surf - tsel) < tol || fabs((double)dt) < tol;
        if (dt > (R)0) {
          const double t_end = exp((double)log_t - (double)sigma * (double)dt);
          frag |= fabs(t_end - (double)depth_q) < 1e-5 || fabs(t_end - (double)threshold) < 1e-5 * threshold;
        }
      }
      if (dt > (R)0) {
        R dlt = -sigma * dt;
        R alpha = (R)1 - FN(rexp)(dlt);
        rgb = FN(add)(rgb, FN(scl)(pf.rgb, alpha * trans));
        nacc = FN(add)(nacc, FN(scl)(FN(ld3)(ns), alpha * trans));
        log_t += dlt;
        if (qstep < 0 && FN(rexp)(log_t) < (R)depth_q) {
          depth = tn + FN(rlog)(trans / (R)depth_q) / sigma;
          qstep = step;
          qsig = sigma;
        }
      }
      prim = next;
      step++;
    }
    out->color[3 * px + 0] = rgb.x;
    out->color[3 * px + 1] = rgb.y;
    out->color[3 * px + 2] = rgb.z;
    if (out->log_t) out->log_t[px] = log_t;
    if (out->normal) {
      out->normal[3 * px + 0] = nacc.x;
      out->normal[3 * px + 1] = nacc.y;
      out->normal[3 * px + 2] = nacc.z;
    }
    if (out->depth) out->depth[px] = depth;
    if (out->q_step) out->q_step[px] = qstep;
    if (out->q_sigma) out->q_sigma[px] = qsig;
    if (out->n_cells) out->n_cells[px] = step;
    if (out->fragile) out->fragile[px] = (uint8_t)frag;
  }
}

/* Exact adjoint of orc_trace_fwd, front to back: the suffix sums a cell's
 * optical depth u = sigma dt acts on are the saved final colour/normal minus
 * the running prefix, so no cell stack is kept:
 *   dL/du_k = g_C.(c_k T_{k+1} - S^C_k) + g_N.(n_k T_{k+1} - S^N_k) + g_O T_end
 *             - [k before the depth cell] g_D / sigma*,
 * plus the depth cell's own terms. Face gradients use the full-precision
 * spheres; texel gradients follow plane_intersection_bwd without cut-offs. */
typedef struct {
  const R* color;
  const R* normal;
  const R* log_t;
  const int32_t* q_step;
  const R* q_sigma;
  const float* g_color;
  const float* g_normal;
  const float* g_depth;
  const float* g_opacity;
} FN(trace_bwd_in);

void FN(orc_trace_bwd)(const pf_camera* cam, const pf_scene* sc, const int32_t* start_sites, int start_site,
                       float threshold, float depth_q, int max_steps, const FN(trace_bwd_in) * in,
                       FN(bwd_out) * out) {
  const int H = cam->height, W = cam->width, T = sc->n_texels;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t px = 0; px < (int64_t)H * W; px++) {
    const int i = (int)(px / W), j = (int)(px % W);
    V3 o, d;
    FN(pixel_ray)(cam, i, j, &o, &d);
    const V3 gC = FN(ld3)(in->g_color + 3 * px), gN = FN(ld3)(in->g_normal + 3 * px);
    const R gD = (R)in->g_depth[px], gO = (R)in->g_opacity[px];
    const V3 Cf = FN(mk)(in->color[3 * px], in->color[3 * px + 1], in->color[3 * px + 2]);
    const V3 Nf = FN(mk)(in->normal[3 * px], in->normal[3 * px + 1], in->normal[3 * px + 2]);
    const R Tend = FN(rexp)(in->log_t[px]);
    const int qstep = in->q_step[px];
    const R qsig = in->q_sigma[px];
    int prim = start_sites ? start_sites[px] : start_site;
    R log_t = 0;
    int step = 0;
    V3 Cp = FN(mk)(0, 0, 0), Np = FN(mk)(0, 0, 0);
    while (1) {
      R trans = FN(rexp)(log_t);
      if (trans < (R)threshold || prim == 0x7FFFFFFF || prim < 0 || step >= max_steps) break;
      const float* sp = sc->spheres + 4 * (size_t)prim;
      const float* ns = sc->nsigmas + 4 * (size_t)prim;
      V3 c = FN(ld3)(sp);
      R r = (R)sp[3];
      FN(sphere_hit) h = FN(ray_sphere)(o, d, c, r);
      int hit = h.hit;
      if (FN(len)(FN(sub)(c, o)) < (R)4 * r) hit = 0;
      R tn = h.t_near, tf = h.t_far;
      int tn_id = tn < (R)1e-8 ? -2 : -1, tf_id = tf < (R)1e-8 ? -2 : -1;
      const int a0 = sc->adjacency_offsets[prim], a1 = sc->adjacency_offsets[prim + 1];
      const int n_adj = a1 - a0;
      int next = 0x7FFFFFFF;
      R pt_far = (R)1e10;
      for (int e = a0; e < a1; e++) {
        V3 n;
        R pm;
        FN(trace_face)(sc, e, c, r, &n, &pm);
        R dp = FN(dot)(d, n);
        R t = (pm - FN(dot)(o, n)) / dp;
        if (dp >= (R)0 && t < pt_far) {
          next = sc->adjacency[e];
          pt_far = t;
        }
        if (dp >= (R)0 && t < tf) tf_id = e - a0;
        tf = dp >= (R)0 ? FN(rmin)(t, tf) : tf;
        if (dp < (R)0 && t > tn) tn_id = e - a0;
        tn = dp < (R)0 ? FN(rmax)(t, tn) : tn;
      }
      R sigma = (R)ns[3];
      if (!hit || tn > tf || sigma < (R)1e-3) {
        prim = next;
        step++;
        continue;
      }
      const R tn0 = tn;
      const int tn0_id = tn_id, tf0_id = tf_id;
      const float* tsites = sc->texel_sites + (size_t)prim * T * 3;
      const float* trgbh = sc->texel_rgbh + (size_t)prim * T * 4;
      V3 pn = FN(ld3)(ns);
      FN(plane_fwd_out) pf = FN(plane_fwd)(o, d, tn, c, pn, r, 1, tsites, trgbh, T);
      if (pf.dp >= (R)0 && pf.t_surf < tf) tf_id = n_adj;
      tf = pf.dp >= (R)0 ? FN(rmin)(pf.t_surf, tf) : tf;
      if (pf.dp < (R)0 && pf.t_surf > tn) tn_id = n_adj;
      tn = pf.dp < (R)0 ? FN(rmax)(pf.t_surf, tn) : tn;
      R dt = tf - tn;
      if (dt > (R)0) {
        R u = sigma * dt;
        R e_u = FN(rexp)(-u);
        R alpha = (R)1 - e_u;
        R Tn = trans * e_u;
        V3 Ck = FN(scl)(pf.rgb, alpha * trans), Nk = FN(scl)(pn, alpha * trans);
        Cp = FN(add)(Cp, Ck);
        Np = FN(add)(Np, Nk);
        V3 SC = FN(sub)(Cf, Cp), SN = FN(sub)(Nf, Np);
        R dLdu = FN(dot)(gC, FN(sub)(FN(scl)(pf.rgb, Tn), SC)) + FN(dot)(gN, FN(sub)(FN(scl)(pn, Tn), SN)) +
                 gO * Tend;
        R dLdsig = 0, dLdtn = 0, dLdtf = 0;
        if (qstep >= 0 && step < qstep) dLdu += -gD / qsig;
        if (step == qstep) {
          dLdtn += gD;
          dLdsig += -gD * FN(rlog)(trans / (R)depth_q) / (sigma * sigma);
        }
        dLdsig += dLdu * dt;
        dLdtf += dLdu * sigma;
        dLdtn -= dLdu * sigma;
        V3 dLdcol = FN(scl)(gC, alpha * trans);
        V3 dLdnorm = FN(scl)(gN, alpha * trans);
        R dLdts = (tn_id == n_adj) ? dLdtn : ((tf_id == n_adj) ? dLdtf : (R)0);
        V3 pdc, pdn;
        R dtq;
        FN(plane_bwd)(o, d, tn0, c, pn, r, 1, tsites, trgbh, &pf, dLdcol, dLdts, T, 1,
                      out->grad_texel_sites + (size_t)prim * T * 3, out->grad_texel_rgb + (size_t)prim * T * 3,
                      out->grad_texel_height + (size_t)prim * T, &pdc, &pdn, &dtq);
        V3 dLdc = pdc;
        dLdnorm = FN(add)(dLdnorm, pdn);
        R dLdr = 0;
        if (tn_id != n_adj) dtq += dLdtn;
        for (int side = 0; side < 2; side++) {
          int tid = side == 0 ? tn0_id : tf0_id;
          R g = side == 0 ? dtq : dLdtf;
          if (side == 1 && tf_id == n_adj) continue;
          if (tid == -1) {
            FN(sphere_grad) sg = FN(ray_sphere_bwd)(o, d, c, r);
            dLdc = FN(add)(dLdc, FN(scl)(side == 0 ? sg.dtn_dc : sg.dtf_dc, g));
            dLdr += g * (side == 0 ? sg.dtn_dr : sg.dtf_dr);
          } else if (tid >= 0 && tid < n_adj) {
            int adj = sc->adjacency[a0 + tid];
            const float* asp = sc->spheres + 4 * (size_t)adj;
            V3 dtdc, dtda;
            R dtdr, dtdra;
            FN(ray_pface_bwd)(o, d, c, r, FN(ld3)(asp), (R)asp[3], &dtdc, &dtdr, &dtda, &dtdra);
            dLdc = FN(add)(dLdc, FN(scl)(dtdc, g));
            dLdr += g * dtdr;
            R* ga = out->grad_spheres + 4 * (size_t)adj;
            FN(atomic_addr)(&ga[0], g * dtda.x);
            FN(atomic_addr)(&ga[1], g * dtda.y);
            FN(atomic_addr)(&ga[2], g * dtda.z);
            FN(atomic_addr)(&ga[3], g * dtdra);
          }
        }
        R* gs = out->grad_spheres + 4 * (size_t)prim;
        FN(atomic_addr)(&gs[0], dLdc.x);
        FN(atomic_addr)(&gs[1], dLdc.y);
        FN(atomic_addr)(&gs[2], dLdc.z);
        FN(atomic_addr)(&gs[3], dLdr);
        R* gq = out->grad_nsigmas + 4 * (size_t)prim;
        FN(atomic_addr)(&gq[0], dLdnorm.x);
        FN(atomic_addr)(&gq[1], dLdnorm.y);
        FN(atomic_addr)(&gq[2], dLdnorm.z);
        FN(atomic_addr)(&gq[3], dLdsig);
        log_t -= u;
      }
      prim = next;
      step++;
    }
  }
}

#undef V3
#undef FN
#undef CAT
#undef CAT_
