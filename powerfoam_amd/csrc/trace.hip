// Ray tracer: per-pixel walk through the power cells, forward and backward.
//
//   RayTracer.benchmark_kernel  raytrace.py:115-241
//   RayTracer.benchmark         raytrace.py:247-309
//
// Each lane walks one pixel's ray from a start cell; the next cell is the
// neighbour whose face the ray leaves through first (argmin t over faces with
// d.n >= 0, the first in list order on ties). The walk -- ray direction, face
// crossings and the choice of the next cell -- is evaluated without FMA
// contraction and with IEEE division and square root, in the operation order
// of the FP32 oracle, so visited cell sequences are bit-exact; shading uses the
// same fast paths as the rasterizer. One wave covers an 8x8 pixel tile so that
// neighbouring rays walk neighbouring cells, and lanes whose rays have ended
// drop out of the loop (the wave runs while any lane walks).
//
// Outputs beyond the reference (which returns colour only): log T, the
// alpha-weighted normal, the depth where T first drops below depth_q, and the
// visited cells. The backward is the exact adjoint computed front to back: a
// cell's optical depth u = sigma dt acts on the suffix sums of colour and
// normal, which are the saved final values minus the running prefix, so the
// backward re-walks the ray instead of storing it.
#include "pf_common.hpp"

namespace pf {

constexpr int TWAVE = 64;
// One wave per workgroup: a workgroup's wave slots and LDS stay taken until
// its last wave ends, and the walks of neighbouring tiles differ in length
// (on the 1M-site fisheye views four-wave workgroups ran the forward in
// 4.9 ms, one-wave ones in 3.4 ms)
constexpr int TWPB = 1;
constexpr int SENTINEL = 0x7FFFFFFF;

struct TraceParams {
  int W, H, tiles_h, n_tiles;
  int n_sites;
  f3 eye, right, up;
  const float* ray_maps;
  int n_texels;
  const float4* spheres;
  const float4* nsigmas;
  const float* texel_sites;
  const float4* texel_rgbh;
  const int32_t* adjacency;
  const int32_t* adj_off;
  const uint16_t* adj_diff;
  const float4* planes;
  int start_site;
  const int32_t* start_sites;
  float thr, depth_q;
  int max_steps;
  // forward outputs (optional except colour)
  float* color;
  float* log_t;
  float* normal;
  float* depth;
  int32_t* q_step;
  float* q_sigma;
  int32_t* cells;
  int K;
  int32_t* n_cells;
  int32_t* tile_steps;
  const int32_t* tile_order;
  // backward inputs / outputs
  const float* s_color;
  const float* s_normal;
  const float* s_log_t;
  const int32_t* s_q_step;
  const float* s_q_sigma;
  const float* g_color;
  const float* g_normal;
  const float* g_depth;
  const float* g_opacity;
  float* g_spheres;
  float* g_nsigmas;
  float* g_tsites;
  float* g_trgb;
  float* g_theight;
};

// camera.py:19-24 / ray_maps, in the FP32 oracle's operation order
__device__ __forceinline__ void walk_ray(const TraceParams& p, int i, int j, f3& o, f3& d) {
#pragma clang fp contract(off)
  f3 rd;
  if (!p.ray_maps) {
    const f3 cr{p.up.y * p.right.z - p.up.z * p.right.y, p.up.z * p.right.x - p.up.x * p.right.z,
                p.up.x * p.right.y - p.up.y * p.right.x};
    const float l = sqrtf(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
    const f3 fw = l > 0.0f ? f3{cr.x / l, cr.y / l, cr.z / l} : f3{0.0f, 0.0f, 0.0f};
    const float x = 2.0f * (float)j / ((float)p.W - 1.0f) - 1.0f;
    const float y = 1.0f - 2.0f * (float)i / ((float)p.H - 1.0f);
    rd = f3{p.right.x * x + p.up.x * y + fw.x, p.right.y * x + p.up.y * y + fw.y, p.right.z * x + p.up.z * y + fw.z};
    o = p.eye;
  } else {
    const float* rm = p.ray_maps + ((size_t)i * p.W + j) * 6;
    o = f3{rm[0], rm[1], rm[2]};
    rd = f3{rm[3], rm[4], rm[5]};
  }
  const float len = sqrtf(rd.x * rd.x + rd.y * rd.y + rd.z * rd.z);
  d = f3{rd.x / len, rd.y / len, rd.z / len};
}

// Where the walk takes a face's plane (n, pm) from: the neighbour's sphere
// (FM_SPHERES, exact), the binary16 (c_adj - c, pm_adj - pm) record
// (FM_HALF, the reference's benchmark encoding), or the exact plane of the
// directed edge written beforehand by trace_planes_kernel (FM_PLANES: the
// same FP32 values as FM_SPHERES, read as one contiguous record per face).
constexpr int FM_SPHERES = 0, FM_HALF = 1, FM_PLANES = 2;

// The faces of a cell (centre c, radius r) four at a time. A group's records
// (neighbour ids and spheres, binary16 records or planes) are all requested
// before any is used, and the ids of the next group are requested while this
// group's spheres load, so a lane waits about one memory latency per group
// instead of two per face. Slots past the cell's last face repeat that face:
// a repeated face moves no strict comparison of the walk.
struct FaceQuad {
  int id[4];     // neighbour ids (FM_PLANES: face indices, the id read for the winner only)
  f3 n[4];
  float pm[4];
  float4 nb[4];  // neighbour spheres (FM_SPHERES only)
};

__device__ __forceinline__ void quad_ids(const TraceParams& p, int e0, int a1, int id[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) id[k] = p.adjacency[min(e0 + k, a1 - 1)];
}

// group e0 .. e0 + 3 (e0 < a1) from its ids; ids become those of the group at
// e0 + 4 when there is one (FM_PLANES reads no ids)
template <int FM>
__device__ __forceinline__ void quad_faces(const TraceParams& p, int e0, int a1, f3 c, float r, int ids[4],
                                           FaceQuad& q) {
#pragma clang fp contract(off)
  if (FM == FM_PLANES) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = min(e0 + k, a1 - 1);
      const float4 pl = p.planes[e];
      q.id[k] = e;
      q.n[k] = f3{pl.x, pl.y, pl.z};
      q.pm[k] = pl.w;
      q.nb[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) q.id[k] = ids[k];
  if (FM == FM_HALF) {
    uint2 raw[4];
#pragma unroll
    for (int k = 0; k < 4; k++) raw[k] = *reinterpret_cast<const uint2*>(p.adj_diff + 4 * (int64_t)min(e0 + k, a1 - 1));
    if (e0 + 4 < a1) quad_ids(p, e0 + 4, a1, ids);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      q.n[k] = f3{half_bits_to_float((uint16_t)(raw[k].x & 0xffff)), half_bits_to_float((uint16_t)(raw[k].x >> 16)),
                  half_bits_to_float((uint16_t)(raw[k].y & 0xffff))};
      q.pm[k] = half_bits_to_float((uint16_t)(raw[k].y >> 16));
      q.nb[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) q.nb[k] = p.spheres[q.id[k]];
    if (e0 + 4 < a1) quad_ids(p, e0 + 4, a1, ids);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const float4 a = q.nb[k];
      q.n[k] = f3{a.x - c.x, a.y - c.y, a.z - c.z};
      q.pm[k] = 0.5f * ((a.x * a.x + a.y * a.y + a.z * a.z) - a.w * a.w) -
                0.5f * ((c.x * c.x + c.y * c.y + c.z * c.z) - r * r);
    }
  }
}

// Crossing of face (n, pm): t = (pm - o.n) / (d.n), exactly as the oracle.
__device__ __forceinline__ float walk_cross(f3 o, f3 d, f3 n, float pm, float& dp) {
#pragma clang fp contract(off)
  dp = d.x * n.x + d.y * n.y + d.z * n.z;
  const float on = o.x * n.x + o.y * n.y + o.z * n.z;
  return (pm - on) / dp;
}

// Texel colour at the cell entry (plane_intersection_fwd_local,
// raytrace.py:42-113), with the colour-stage weights returned.
struct TexelEval {
  float ts0, ts1, height, ws0, ws, dpn, cn;
  f3 col;
  bool heights;
};

// The 3T texel-site floats of a site in 16-byte pieces (gfx950 reads
// unaligned global data); a piece that would run past the array is read
// float by float.
__device__ __forceinline__ void load_texel_sites(const TraceParams& p, int site, int T, float* v) {
  const int n = 3 * T;
  const size_t base = (size_t)site * n, total = (size_t)p.n_sites * n;
#pragma unroll
  for (int j = 0; j < 3 * PF_MAX_TEXELS; j += 4) {
    if (j < n) {
      if (base + j + 4 <= total) {
        float4 q;
        __builtin_memcpy(&q, p.texel_sites + base + j, sizeof(q));
        v[j] = q.x;
        v[j + 1] = q.y;
        v[j + 2] = q.z;
        v[j + 3] = q.w;
      } else {
#pragma unroll
        for (int k = 0; k < 4; k++) v[j + k] = j + k < n ? p.texel_sites[base + j + k] : 0.0f;
      }
    }
  }
}

// A site's texel record in registers: site(i) = texel site i (from the 3T
// site floats), rgbh(i) = its (rgb, height).
struct RegRec {
  const float* tsv;
  const float4* cl;
  __device__ __forceinline__ f3 site(int i) const { return f3{tsv[3 * i], tsv[3 * i + 1], tsv[3 * i + 2]}; }
  __device__ __forceinline__ float4 rgbh(int i) const { return cl[i]; }
};
// The texel stage (plane_intersection_fwd_local) from a texel record.
template <typename Rec>
__device__ __forceinline__ TexelEval texel_eval_rec(int T, const Rec& rec, f3 c, float r, f3 pn, f3 o, f3 d, float tn,
                                                    float* w1) {
  TexelEval te;
  const f3 co = c - o;
  te.dpn = dot(pn, d);
  te.cn = dot(co, pn);
  te.ts0 = acc_div(te.cn, te.dpn);
  const float kexp = -10.0f * fast_rcp(r * r);
  bool nz = false;
#pragma unroll
  for (int i = 0; i < PF_MAX_TEXELS; i++)
    if (i < T) nz = nz || rec.rgbh(i).w != 0.0f;
  te.heights = nz;
  te.height = 0.0f;
  te.ws0 = 1.0f;
  if (nz) {
    const float tq0 = te.dpn >= 0.0f ? tn : sel_max(tn, te.ts0);
    const f3 x0 = d * tq0 - co;
    float hsum = 0.0f, s0 = 0.0f;
#pragma unroll
    for (int i = 0; i < PF_MAX_TEXELS; i++) {
      if (i < T) {
        const f3 ts = rec.site(i);
        const f3 df{x0.x - (ts.x - c.x), x0.y - (ts.y - c.y), x0.z - (ts.z - c.z)};
        const float w = fast_exp(dot(df, df) * kexp);
        hsum += w * rec.rgbh(i).w;
        s0 += w;
      }
    }
    te.ws0 = sel_max(s0, 1e-20f);
    te.height = hsum * fast_rcp(te.ws0);
    te.ts1 = acc_div(te.cn + te.height, te.dpn);
  } else {
    te.ts1 = te.ts0;
  }
  const float tq = te.dpn >= 0.0f ? tn : sel_max(tn, te.ts1);
  const f3 x1 = d * tq - co;
  f3 rs{0.0f, 0.0f, 0.0f};
  float s1 = 0.0f;
#pragma unroll
  for (int i = 0; i < PF_MAX_TEXELS; i++) {
    if (i < T) {
      const f3 ts = rec.site(i);
      const f3 df{x1.x - (ts.x - c.x), x1.y - (ts.y - c.y), x1.z - (ts.z - c.z)};
      const float w = fast_exp(dot(df, df) * kexp);
      if (w1) w1[i * TWAVE] = w;  // a lane's column of a [texel][lane] LDS array
      const float4 cl = rec.rgbh(i);
      rs = rs + f3{cl.x, cl.y, cl.z} * w;
      s1 += w;
    }
  }
  te.ws = sel_max(s1, 1e-20f);
  if (!nz) te.ws0 = te.ws;
  te.col = rs * fast_rcp(te.ws);
  return te;
}

__device__ __forceinline__ TexelEval texel_eval(const TraceParams& p, int site, f3 c, float r, f3 pn, f3 o, f3 d,
                                                float tn, float* w1) {
  const int T = p.n_texels;
  float4 cl[PF_MAX_TEXELS];
  float tsv[3 * PF_MAX_TEXELS];
  load_texel_sites(p, site, T, tsv);
#pragma unroll
  for (int i = 0; i < PF_MAX_TEXELS; i++)
    if (i < T) cl[i] = p.texel_rgbh[(size_t)site * T + i];
  return texel_eval_rec(T, RegRec{tsv, cl}, c, r, pn, o, d, tn, w1);
}

__device__ __forceinline__ bool trace_lane(const TraceParams& p, int tile, int lane, int& i, int& j) {
  const int ti = tile % p.tiles_h, tj = tile / p.tiles_h;
  i = ti * 8 + (lane & 7);
  j = tj * 8 + (lane >> 3);
  return i < p.H && j < p.W;
}

// COMPACT: a wave walks the rays of group_tiles neighbouring tiles; a lane
// whose ray ends writes its outputs and takes the group's next unstarted ray
// (slot = the wave's counter in LDS + the lane's rank among the lanes that
// ended in the same step, from a ballot), so lanes stay busy until the group
// runs out of rays instead of idling until the tile's longest ray ends.
//
// Without compaction a wave walks one tile, tiles in index order (workgroups
// dealt round-robin to the XCDs, which spreads the costly tiles of the view's
// centre over all eight; contiguous runs per XCD measured 20 % slower), and
// the wave's longest walk is written to tile_steps[tile] when asked (the
// backward's longest-first order).
template <bool COMPACT, int FM>
__global__ void __launch_bounds__(TWAVE* TWPB) trace_fwd_kernel(TraceParams p, int group_tiles) {
  __shared__ int s_next[TWPB];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int group = blockIdx.x * TWPB + wv;
  if (group * group_tiles >= p.n_tiles) return;
  const int n_slots = group_tiles * TWAVE;
  int slot = lane;
  // the tile's first ray (lane 0) always lies in the image, so lane 0 reaches
  // the end of the kernel
  if (COMPACT || p.tile_steps) {
    if (lane == 0) s_next[wv] = COMPACT ? TWAVE : 0;
    wave_sync();
  }
  int i, j;
  // a slot outside the image takes the next one (compaction) or idles
  while (!(group * group_tiles + slot / TWAVE < p.n_tiles &&
           trace_lane(p, group * group_tiles + slot / TWAVE, slot % TWAVE, i, j))) {
    if (!COMPACT) return;
    const uint64_t m = ballot(true);
    const int base = s_next[wv];
    slot = base + __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    wave_sync();
    if (lane == __ffsll((long long)m) - 1) s_next[wv] = base + __popcll(m);
    wave_sync();
    if (slot >= n_slots) return;
  }
  size_t px = (size_t)i * p.W + j;
  f3 o, d;
  walk_ray(p, i, j, o, d);
  int prim = p.start_sites ? p.start_sites[px] : p.start_site;
  // the sphere of the cell the walk enters, when the face loop read it
  float4 csph = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  bool carried = false;
  float log_t = 0.0f, depth = -1.0f, qsig = 0.0f;
  int qstep = -1, step = 0;
  f3 rgb{0.0f, 0.0f, 0.0f}, nacc{0.0f, 0.0f, 0.0f};
  while (true) {
    const float trans = fast_exp(log_t);
    if (trans < p.thr || prim == SENTINEL || prim < 0 || step >= p.max_steps) {
      if (!COMPACT) break;
      p.color[3 * px + 0] = rgb.x;
      p.color[3 * px + 1] = rgb.y;
      p.color[3 * px + 2] = rgb.z;
      if (p.log_t) p.log_t[px] = log_t;
      if (p.normal) {
        p.normal[3 * px + 0] = nacc.x;
        p.normal[3 * px + 1] = nacc.y;
        p.normal[3 * px + 2] = nacc.z;
      }
      if (p.depth) p.depth[px] = depth;
      if (p.q_step) p.q_step[px] = qstep;
      if (p.q_sigma) p.q_sigma[px] = qsig;
      if (p.n_cells) p.n_cells[px] = step;
      // the next unstarted ray of the group (skipping slots outside the image)
      bool got = false;
      while (true) {
        const uint64_t m = ballot(true);
        const int base = s_next[wv];
        slot = base + __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        wave_sync();
        if (lane == __ffsll((long long)m) - 1) s_next[wv] = base + __popcll(m);
        wave_sync();
        if (slot >= n_slots) break;
        const int tl = group * group_tiles + slot / TWAVE;
        if (tl < p.n_tiles && trace_lane(p, tl, slot % TWAVE, i, j)) {
          got = true;
          break;
        }
      }
      if (!got) break;
      px = (size_t)i * p.W + j;
      walk_ray(p, i, j, o, d);
      prim = p.start_sites ? p.start_sites[px] : p.start_site;
      carried = false;
      log_t = 0.0f;
      depth = -1.0f;
      qsig = 0.0f;
      qstep = -1;
      step = 0;
      rgb = f3{0.0f, 0.0f, 0.0f};
      nacc = f3{0.0f, 0.0f, 0.0f};
      continue;
    }
    if (p.cells && step < p.K) p.cells[px * p.K + step] = prim;
    const float4 sph = carried ? csph : p.spheres[prim];
    const int a0 = p.adj_off[prim], a1 = p.adj_off[prim + 1];
    const float4 ns = p.nsigmas[prim];
    const f3 c{sph.x, sph.y, sph.z};
    const float r = sph.w;
    float tn, tf;
    const f3 oc = o - c;
    bool hit = sphere_hit(oc, d, r, tn, tf);
    if (__builtin_sqrtf(dot(oc, oc)) < 4.0f * r) hit = false;
    int next = SENTINEL;
    float pt_far = 1e10f;
    int ids[4];
    if (FM != FM_PLANES && a0 < a1) quad_ids(p, a0, a1, ids);
    for (int e0 = a0; e0 < a1; e0 += 4) {
      FaceQuad q;
      quad_faces<FM>(p, e0, a1, c, r, ids, q);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        float dp;
        const float t = walk_cross(o, d, q.n[k], q.pm[k], dp);
        if (dp >= 0.0f && t < pt_far) {
          next = q.id[k];
          csph = q.nb[k];
          pt_far = t;
        }
        tf = (dp >= 0.0f && t < tf) ? t : tf;
        tn = (dp < 0.0f && t > tn) ? t : tn;
      }
    }
    if (FM == FM_PLANES && next != SENTINEL) next = p.adjacency[next];
    carried = FM == FM_SPHERES;
    const float sigma = ns.w;
    if (hit && !(tn > tf) && !(sigma < 1e-3f)) {
      const f3 pn{ns.x, ns.y, ns.z};
      const TexelEval te = texel_eval(p, prim, c, r, pn, o, d, tn, nullptr);
      const float tf2 = (te.dpn >= 0.0f && te.ts1 < tf) ? te.ts1 : tf;
      const float tn2 = (te.dpn < 0.0f && te.ts1 > tn) ? te.ts1 : tn;
      const float dt = tf2 - tn2;
      if (dt > 0.0f) {
        const float dlt = -sigma * dt;
        const float at = (1.0f - fast_exp(dlt)) * trans;
        rgb = rgb + te.col * at;
        nacc = nacc + pn * at;
        log_t += dlt;
        if (qstep < 0 && fast_exp(log_t) < p.depth_q) {
          depth = tn2 + __logf(trans * fast_rcp(p.depth_q)) * fast_rcp(sigma);
          qstep = step;
          qsig = sigma;
        }
      }
    }
    prim = next;
    step++;
  }
  p.color[3 * px + 0] = rgb.x;
  p.color[3 * px + 1] = rgb.y;
  p.color[3 * px + 2] = rgb.z;
  if (p.log_t) p.log_t[px] = log_t;
  if (p.normal) {
    p.normal[3 * px + 0] = nacc.x;
    p.normal[3 * px + 1] = nacc.y;
    p.normal[3 * px + 2] = nacc.z;
  }
  if (p.depth) p.depth[px] = depth;
  if (p.q_step) p.q_step[px] = qstep;
  if (p.q_sigma) p.q_sigma[px] = qsig;
  if (p.n_cells) p.n_cells[px] = step;
  if (!COMPACT && p.tile_steps) {
    __hip_atomic_fetch_max((__attribute__((address_space(3))) int*)&s_next[wv], step, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WAVEFRONT);
    wave_sync();
    if (lane == 0) p.tile_steps[group] = s_next[wv];
  }
}

// Gradient cache of a wave. The 64 rays of a wave walk mostly the same
// cells (on a 1M-site fisheye view a wave visits ~110 distinct cells for
// 64 x 50 visits), so a visit adds its site's gradients (texel colours and
// positions, heights, sphere, normal and density: 7T + 8 values) into an LDS
// slot of that site, and a slot goes out with one atomic per value when
// another site claims it or the wave ends. Slot = site mod GC_SLOTS; a lane
// whose slot another lane's site won this step adds to global memory
// directly. The sphere gradients a visit sends to the neighbours across its
// entry and exit faces join the neighbour's slot the same way when the slot
// holds that neighbour. Lanes walk a site's texels (and the four neighbour
// values) in an order rotated by lane, so lanes on one site add to different
// words of a slot at a time (an LDS atomic serialises the lanes that hit one
// word). Only the summation order of the gradients changes.
constexpr int GC_SLOTS = 32;
constexpr int GC_VALS = 7 * PF_MAX_TEXELS + 8;

struct GradCache {
  int tag[GC_SLOTS];
  float v[GC_SLOTS][GC_VALS];
};

// address of gradient value k (< 7T + 8) of site s
__device__ __forceinline__ float* grad_addr(const TraceParams& p, int s, int k, int T) {
  if (k < 3 * T) return p.g_trgb + (size_t)s * 3 * T + k;
  k -= 3 * T;
  if (k < 3 * T) return p.g_tsites + (size_t)s * 3 * T + k;
  k -= 3 * T;
  if (k < T) return p.g_theight + (size_t)s * T + k;
  k -= T;
  if (k < 4) return p.g_spheres + 4 * (size_t)s + k;
  return p.g_nsigmas + 4 * (size_t)s + (k - 4);
}

// all 64 lanes: slot q's values to global memory (7T + 8 <= 64 values)
__device__ __forceinline__ void gc_flush(const TraceParams& p, const GradCache& gc, int q, int lane, int T) {
  const int s = gc.tag[q];
  if (s >= 0 && lane < 7 * T + 8) {
    const float v = gc.v[q][lane];
    if (v != 0.0f) global_add(grad_addr(p, s, lane, T), v);
  }
}

template <int FM>
__global__ void __launch_bounds__(TWAVE* TWPB) __attribute__((amdgpu_waves_per_eu(3)))
trace_bwd_kernel(TraceParams p) {
  __shared__ GradCache s_gc[TWPB];
  // per lane, the colour-stage texel weights of its visit ([texel][lane]),
  // written by texel_eval
  __shared__ float s_wn_all[TWPB][PF_MAX_TEXELS * TWAVE];
  float* s_wn = s_wn_all[threadIdx.x >> 6];
  // slot b * TWPB + wave takes tile tile_order[slot], or tile slot
  const int slot = blockIdx.x * TWPB + (threadIdx.x >> 6);
  if (slot >= p.n_tiles) return;
  const int tile = p.tile_order ? p.tile_order[slot] : slot;
  const int lane = threadIdx.x & 63;
  GradCache& gc = s_gc[threadIdx.x >> 6];
  if (lane < GC_SLOTS) gc.tag[lane] = -1;
  int i, j;
  // lanes past the image edge take part in the wave's cache bookkeeping only
  bool live = trace_lane(p, tile, lane, i, j);
  const size_t px = live ? (size_t)i * p.W + j : 0;
  f3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 1.0f};
  if (live) walk_ray(p, i, j, o, d);
  const int T = p.n_texels;
  const f3 gC = live ? ld3(p.g_color + 3 * px) : f3{0.0f, 0.0f, 0.0f};
  const f3 gN = live ? ld3(p.g_normal + 3 * px) : f3{0.0f, 0.0f, 0.0f};
  const float gD = (live && p.g_depth) ? p.g_depth[px] : 0.0f;
  const float gO = live ? p.g_opacity[px] : 0.0f;
  const f3 Cf = live ? ld3(p.s_color + 3 * px) : f3{0.0f, 0.0f, 0.0f};
  const f3 Nf = live ? ld3(p.s_normal + 3 * px) : f3{0.0f, 0.0f, 0.0f};
  const float Tend = live ? fast_exp(p.s_log_t[px]) : 0.0f;
  const int qstep = live ? p.s_q_step[px] : -1;
  const float qsig = live ? p.s_q_sigma[px] : 1.0f;
  int prim = live ? (p.start_sites ? p.start_sites[px] : p.start_site) : SENTINEL;
  float4 csph = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  bool carried = false;
  float log_t = 0.0f;
  int step = 0;
  f3 Cp{0.0f, 0.0f, 0.0f}, Np{0.0f, 0.0f, 0.0f};
  wave_sync();
  while (true) {
    const float trans = fast_exp(log_t);
    if (live && (trans < p.thr || prim == SENTINEL || prim < 0 || step >= p.max_steps)) live = false;
    if (ballot(live) == 0) break;
    float4 sph = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (live) sph = carried ? csph : p.spheres[prim];
    const int a0 = live ? p.adj_off[prim] : 0, a1 = live ? p.adj_off[prim + 1] : 0;
    float4 ns = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (live) ns = p.nsigmas[prim];
    const f3 c{sph.x, sph.y, sph.z};
    const float r = sph.w;
    float tn = 0.0f, tf = 0.0f;
    const f3 oc = o - c;
    bool hit = live && sphere_hit(oc, d, r, tn, tf);
    if (__builtin_sqrtf(dot(oc, oc)) < 4.0f * r) hit = false;
    int tn_id = tn < 1e-8f ? -2 : -1, tf_id = tf < 1e-8f ? -2 : -1;
    int next = SENTINEL;
    float pt_far = 1e10f;
    const int n_adj = a1 - a0;
    int ids[4];
    if (FM != FM_PLANES && a0 < a1) quad_ids(p, a0, a1, ids);
    for (int e0 = a0; e0 < a1; e0 += 4) {
      FaceQuad q;
      quad_faces<FM>(p, e0, a1, c, r, ids, q);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        float dp;
        const float t = walk_cross(o, d, q.n[k], q.pm[k], dp);
        if (dp >= 0.0f && t < pt_far) {
          next = q.id[k];
          csph = q.nb[k];
          pt_far = t;
        }
        if (dp >= 0.0f && t < tf) {
          tf_id = e0 + k - a0;
          tf = t;
        }
        if (dp < 0.0f && t > tn) {
          tn_id = e0 + k - a0;
          tn = t;
        }
      }
    }
    if (FM == FM_PLANES && next != SENTINEL) next = p.adjacency[next];
    carried = FM == FM_SPHERES;
    const float sigma = ns.w;
    const f3 pn{ns.x, ns.y, ns.z};
    const f3 co = -oc;
    const float tn0 = tn;
    const int tn0_id = tn_id, tf0_id = tf_id;
    TexelEval te{};
    float dt = -1.0f;
    if (live && hit && !(tn > tf) && !(sigma < 1e-3f)) {
      te = texel_eval(p, prim, c, r, pn, o, d, tn, s_wn + lane);
      if (te.dpn >= 0.0f && te.ts1 < tf) {
        tf_id = n_adj;
        tf = te.ts1;
      }
      if (te.dpn < 0.0f && te.ts1 > tn) {
        tn_id = n_adj;
        tn = te.ts1;
      }
      dt = tf - tn;
    }
    const bool act = dt > 0.0f;
    // the cache slot of this visit's site
    const int slot = prim & (GC_SLOTS - 1);
    bool in_lds = false;
    wave_sync();  // the last step's cache adds land before any flush reads them
    {
      bool pending = act;
      if (pending && gc.tag[slot] == prim) {
        in_lds = true;
        pending = false;
      }
      uint64_t pm;
      while ((pm = ballot(pending)) != 0) {
        const int leader = __builtin_ctzll(pm);
        const int ls = readlane(slot, leader), lp = readlane(prim, leader);
        // a slot that holds a site visited this step stays; its rivals go direct
        const bool taken = ballot(in_lds && slot == ls) != 0;
        if (!taken) {
          gc_flush(p, gc, ls, lane, T);
          wave_sync();
          if (lane < GC_VALS) gc.v[ls][lane] = 0.0f;
          if (lane == 0) gc.tag[ls] = lp;
          wave_sync();
        }
        if (pending && slot == ls) {
          in_lds = !taken && prim == lp;
          pending = false;
        }
      }
    }
    float* cache = gc.v[slot];
    // value k of this visit's site gradient: into the cache, or to global memory
    auto acc = [&](int k, float v) {
      if (in_lds)
        lds_add(cache + k, v);
      else
        global_add(grad_addr(p, prim, k, T), v);
    };
    if (act) {
        const float u = sigma * dt;
        const float e_u = fast_exp(-u);
        const float alpha = 1.0f - e_u;
        const float Tn = trans * e_u;
        Cp = Cp + te.col * (alpha * trans);
        Np = Np + pn * (alpha * trans);
        const f3 SC = Cf - Cp, SN = Nf - Np;
        float dLdu = dot(gC, te.col * Tn - SC) + dot(gN, pn * Tn - SN) + gO * Tend;
        float dLdsig = 0.0f, dLdtn = 0.0f, dLdtf = 0.0f;
        if (qstep >= 0 && step < qstep) dLdu += -gD / qsig;
        if (step == qstep) {
          dLdtn += gD;
          dLdsig += -gD * __logf(trans * fast_rcp(p.depth_q)) / (sigma * sigma);
        }
        dLdsig += dLdu * dt;
        dLdtf += dLdu * sigma;
        dLdtn -= dLdu * sigma;
        const f3 dLdcol = gC * (alpha * trans);
        f3 dLdnorm = gN * (alpha * trans);
        const float dLdts_in = (tn_id == n_adj) ? dLdtn : ((tf_id == n_adj) ? dLdtf : 0.0f);
        // plane_intersection_bwd (texture.py:74-186), exact form
        const float kw = 20.0f * fast_rcp(r * r);
        const float kexp = -10.0f * fast_rcp(r * r);
        const float inv_ws = fast_rcp(te.ws), inv_ws0 = fast_rcp(te.ws0);
        float dLdheight = 0.0f;
        f3 pdc{0.0f, 0.0f, 0.0f}, pdn{0.0f, 0.0f, 0.0f};
        const float tq = te.dpn >= 0.0f ? tn0 : sel_max(tn0, te.ts1);
        const f3 x1 = d * tq - co;
        f3 dLdx{0.0f, 0.0f, 0.0f};
        const float gdp_in = dot(dLdcol, te.col);
        // lanes start at texel (lane mod T): lanes on one site then add to T
        // different cache words at a time instead of all to the same one
        const int rot = lane % T;
        for (int kk = 0; kk < T; kk++) {
          const int k = kk + rot < T ? kk + rot : kk + rot - T;
          const float* ts = p.texel_sites + ((size_t)prim * T + k) * 3;
          const float4 cl = p.texel_rgbh[(size_t)prim * T + k];
          const f3 df{x1.x - (ts[0] - c.x), x1.y - (ts[1] - c.y), x1.z - (ts[2] - c.z)};
          // the colour-stage weight texel_eval formed (kept per lane in LDS)
          const float wn = lds_load(s_wn + k * TWAVE + lane) * inv_ws;
          acc(3 * k + 0, dLdcol.x * wn);
          acc(3 * k + 1, dLdcol.y * wn);
          acc(3 * k + 2, dLdcol.z * wn);
          const float gdp = dLdcol.x * cl.x + dLdcol.y * cl.y + dLdcol.z * cl.z;
          const f3 sg = df * (kw * wn * (gdp - gdp_in));
          acc(3 * T + 3 * k + 0, sg.x);
          acc(3 * T + 3 * k + 1, sg.y);
          acc(3 * T + 3 * k + 2, sg.z);
          dLdx = dLdx - sg;
        }
        const float dLdtq = dot(dLdx, d);
        float dtn_q = (te.dpn >= 0.0f || tn0 >= te.ts1) ? dLdtq : 0.0f;
        float dts = (te.dpn < 0.0f && tn0 < te.ts1) ? dLdtq : 0.0f;
        dts += dLdts_in;
        if (dts != 0.0f) {
          const float iden = 1.0f / te.dpn;
          pdc = pdc + pn * (dts * iden);
          pdn = pdn + (co * iden - d * ((te.cn + te.height) / (te.dpn * te.dpn))) * dts;
          dLdheight += dts * iden;
        }
        const float tq0 = te.dpn >= 0.0f ? tn0 : sel_max(tn0, te.ts0);
        const f3 x0 = d * tq0 - co;
        f3 dLdx0{0.0f, 0.0f, 0.0f};
        const float gdp0_in = dLdheight * te.height;
        if (te.heights) {
          for (int kk = 0; kk < T; kk++) {
            const int k = kk + rot < T ? kk + rot : kk + rot - T;
            const float* ts = p.texel_sites + ((size_t)prim * T + k) * 3;
            const float hk = p.texel_rgbh[(size_t)prim * T + k].w;
            const f3 df{x0.x - (ts[0] - c.x), x0.y - (ts[1] - c.y), x0.z - (ts[2] - c.z)};
            const float wn0 = fast_exp(dot(df, df) * kexp) * inv_ws0;
            acc(6 * T + k, wn0 * dLdheight);
            const f3 sg = df * (kw * wn0 * (dLdheight * hk - gdp0_in));
            acc(3 * T + 3 * k + 0, sg.x);
            acc(3 * T + 3 * k + 1, sg.y);
            acc(3 * T + 3 * k + 2, sg.z);
            dLdx0 = dLdx0 - sg;
          }
        } else {
          // all heights zero: x0 is the colour-stage point, ws0 = ws, the
          // weights are the colour stage's, and hk = gdp0_in = 0 leaves no
          // texel-site term
          for (int kk = 0; kk < T; kk++) {
            const int k = kk + rot < T ? kk + rot : kk + rot - T;
            acc(6 * T + k, lds_load(s_wn + k * TWAVE + lane) * inv_ws0 * dLdheight);
          }
        }
        const float dLdtq0 = dot(dLdx0, d);
        dtn_q += (te.dpn >= 0.0f || tn0 >= te.ts0) ? dLdtq0 : 0.0f;
        const float dts0 = (te.dpn < 0.0f && tn0 < te.ts0) ? dLdtq0 : 0.0f;
        if (dts0 != 0.0f) {
          const float iden = 1.0f / te.dpn;
          pdc = pdc + pn * (dts0 * iden);
          pdn = pdn + (co * iden - d * (te.cn / (te.dpn * te.dpn))) * dts0;
        }
        f3 dLdc = pdc;
        dLdnorm = dLdnorm + pdn;
        float dLdr = 0.0f;
        if (tn_id != n_adj) dtn_q += dLdtn;
        for (int side = 0; side < 2; side++) {
          const int tid = side == 0 ? tn0_id : tf0_id;
          const float g = side == 0 ? dtn_q : dLdtf;
          if (side == 1 && tf_id == n_adj) continue;
          if (tid == -1) {
            const f3 cr = cross(oc, d);
            const float sq = 2.0f * __builtin_sqrtf(fmaxf(r * r - dot(cr, cr), 0.0f));
            const float qb = 2.0f * dot(oc, d);
            const float tnear = (-qb - sq) * 0.5f;
            const f3 w = (oc * 2.0f - d * qb) * (1.0f / sq);
            if (side == 0) {
              if (!(tnear < 0.0f)) {
                dLdc = dLdc + (d - w) * g;
                dLdr += g * (-2.0f * r / sq);
              }
            } else {
              dLdc = dLdc + (d + w) * g;
              dLdr += g * (2.0f * r / sq);
            }
          } else if (tid >= 0 && tid < n_adj) {
            const int adj = p.adjacency[a0 + tid];
            const float4 as = p.spheres[adj];
            const f3 a{as.x, as.y, as.z};
            const f3 n = a - c;
            const float num = dot(co, n) + 0.5f * (dot(n, n) + r * r - as.w * as.w);
            const float idp = 1.0f / dot(d, n);
            const f3 qv = d * (num * idp * idp);
            dLdc = dLdc + (oc * idp + qv) * g;
            dLdr += g * (r * idp);
            const f3 ga = ((a - o) * idp - qv) * g;
            // the neighbour is mostly a cell this wave walks too (the one
            // before or after): its sphere gradient joins that site's cache
            // slot when the slot holds it (tags only change at the top of a
            // step), else goes to global memory
            const int aslot = adj & (GC_SLOTS - 1);
            const bool adj_lds = gc.tag[aslot] == adj;
            const float gr = g * (-as.w * idp);
#pragma unroll
            for (int i = 0; i < 4; i++) {
              const int k = (i + lane) & 3;
              const float v = k == 0 ? ga.x : k == 1 ? ga.y : k == 2 ? ga.z : gr;
              if (adj_lds)
                lds_add(gc.v[aslot] + 7 * T + k, v);
              else
                global_add(p.g_spheres + 4 * (size_t)adj + k, v);
            }
          }
        }
        acc(7 * T + 0, dLdc.x);
        acc(7 * T + 1, dLdc.y);
        acc(7 * T + 2, dLdc.z);
        acc(7 * T + 3, dLdr);
        acc(7 * T + 4, dLdnorm.x);
        acc(7 * T + 5, dLdnorm.y);
        acc(7 * T + 6, dLdnorm.z);
        acc(7 * T + 7, dLdsig);
        log_t -= u;
    }
    if (live) {
      prim = next;
      step++;
    }
  }
  wave_sync();
  for (int q = 0; q < GC_SLOTS; q++) gc_flush(p, gc, q, lane, T);
}

static TraceParams make_params(const pf_trace_args* a) {
  TraceParams p{};
  const pf_camera& c = a->cam;
  p.W = c.width;
  p.H = c.height;
  p.tiles_h = tiles_h(c);
  p.n_tiles = tiles_h(c) * tiles_w(c);
  p.eye = f3{c.eye[0], c.eye[1], c.eye[2]};
  p.right = f3{c.right[0], c.right[1], c.right[2]};
  p.up = f3{c.up[0], c.up[1], c.up[2]};
  p.ray_maps = c.ray_maps;
  p.n_texels = a->scene.n_texels;
  p.n_sites = a->scene.n_sites;
  p.spheres = reinterpret_cast<const float4*>(a->scene.spheres);
  p.nsigmas = reinterpret_cast<const float4*>(a->scene.nsigmas);
  p.texel_sites = a->scene.texel_sites;
  p.texel_rgbh = reinterpret_cast<const float4*>(a->scene.texel_rgbh);
  p.adjacency = a->scene.adjacency;
  p.adj_off = a->scene.adjacency_offsets;
  p.adj_diff = a->scene.adjacency_diff;
  p.planes = a->scene.adjacency_diff ? nullptr : reinterpret_cast<const float4*>(a->face_planes);
  p.start_site = a->start_site;
  p.start_sites = a->start_sites;
  p.thr = a->transmittance_threshold;
  p.depth_q = a->depth_quantile;
  p.max_steps = a->max_steps > 0 ? a->max_steps : (1 << 16);
  return p;
}

// ---- FP64 recomputation of the walk on sampled rays ----------------------
// For ray r (pixel pixels[r], generated exactly as the walk generates it) and
// each cell the FP32 walk visited, the cell's interval in FP64: the sphere
// interval (no hit when |o - c| < 4r, as the walk) cut by every face plane,
// entering faces (d.n < 0) raising the entry, leaving faces lowering the exit;
// the FP64 choice of the next cell (argmin exit-face time, first in list order
// on ties) and the gap to the second-smallest exit-face time.
__global__ void __launch_bounds__(256)
    trace_check_f64_kernel(TraceParams p, int n_rays, const int32_t* __restrict__ pixels,
                           const int32_t* __restrict__ cells, const int32_t* __restrict__ n_cells, int K,
                           double* __restrict__ t_enter, double* __restrict__ t_exit, int32_t* __restrict__ next64,
                           double* __restrict__ gap, double* __restrict__ tol) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rays) return;
  const int px = pixels[r];
  f3 of, df;
  walk_ray(p, px / p.W, px % p.W, of, df);
  const double ox = of.x, oy = of.y, oz = of.z, dx = df.x, dy = df.y, dz = df.z;
  const int steps = min(n_cells[r], K);
  for (int k = 0; k < K; k++) {
    const size_t q = (size_t)r * K + k;
    if (k >= steps) {
      t_enter[q] = t_exit[q] = __builtin_nan("");
      next64[q] = -1;
      gap[q] = __builtin_nan("");
      tol[q] = __builtin_nan("");
      continue;
    }
    const int prim = cells[q];
    const float4 s4 = p.spheres[prim];
    const double cx = s4.x, cy = s4.y, cz = s4.z, rr = s4.w;
    const double ocx = ox - cx, ocy = oy - cy, ocz = oz - cz;
    const double b = ocx * dx + ocy * dy + ocz * dz;
    const double crx = ocy * dz - ocz * dy, cry = ocz * dx - ocx * dz, crz = ocx * dy - ocy * dx;
    const double h2 = rr * rr - (crx * crx + cry * cry + crz * crz);
    double tn = 0.0, tf = 0.0;
    bool hit = h2 >= 0.0 && sqrt(ocx * ocx + ocy * ocy + ocz * ocz) >= 4.0 * rr;
    if (hit) {
      const double h = sqrt(h2);
      const double nr = -b - h, fr = -b + h;
      if (nr < 0.0 && fr < 0.0) hit = false;
      tn = nr < 0.0 ? 0.0 : nr;
      tf = fr;
    }
    const double pc = 0.5 * (cx * cx + cy * cy + cz * cz - rr * rr);
    double best = 1e300, second = 1e300, err_best = 0.0, err_second = 0.0;
    int nxt = SENTINEL;
    // rounding bound of the FP32 walk's t = (pm - o.n) / (d.n): unit roundoff
    // times the magnitudes its terms are formed from
    const double eps = 5.960464477539063e-08;
    const double olen = sqrt(ox * ox + oy * oy + oz * oz), clen = sqrt(cx * cx + cy * cy + cz * cz);
    for (int e = p.adj_off[prim]; e < p.adj_off[prim + 1]; e++) {
      const int a = p.adjacency[e];
      const float4 a4 = p.spheres[a];
      const double ax = a4.x, ay = a4.y, az = a4.z, ar = a4.w;
      const double nx = ax - cx, ny = ay - cy, nz = az - cz;
      const double pm = 0.5 * (ax * ax + ay * ay + az * az - ar * ar) - pc;
      const double dp = dx * nx + dy * ny + dz * nz;
      const double t = (pm - (ox * nx + oy * ny + oz * nz)) / dp;
      if (dp >= 0.0) {
        const double alen = sqrt(ax * ax + ay * ay + az * az);
        const double num = fabs(pm) + 0.5 * (alen * alen + ar * ar + clen * clen + rr * rr) +
                           3.0 * (fabs(ox * nx) + fabs(oy * ny) + fabs(oz * nz)) + olen * (alen + clen);
        const double den = 3.0 * (fabs(dx * nx) + fabs(dy * ny) + fabs(dz * nz)) + alen + clen;
        const double err = 4.0 * eps * (num + fabs(t) * den) / fabs(dp);
        if (t < best) {
          second = best;
          err_second = err_best;
          best = t;
          err_best = err;
          nxt = a;
        } else if (t < second) {
          second = t;
          err_second = err;
        }
        tf = t < tf ? t : tf;
      } else {
        tn = t > tn ? t : tn;
      }
    }
    t_enter[q] = hit ? tn : __builtin_nan("");
    t_exit[q] = hit ? tf : __builtin_nan("");
    next64[q] = nxt;
    gap[q] = second - best;
    tol[q] = err_best + err_second;
  }
}

// Exact plane of every directed face: for the face between site s (centre
// c, radius r) and neighbour a = adjacency[e], n = a - c and
// pm = (|a|^2 - r_a^2) / 2 - (|c|^2 - r^2) / 2, formed exactly as the walk
// forms them from the spheres (FP32, no contraction, same operation order),
// so a walk that reads them makes the same choices bit for bit.
// A wave takes 64 consecutive sites and deals their faces to its lanes in
// order (coalesced id reads and plane writes); a lane finds the site of its
// face by binary search over the wave's 65 offsets in LDS.
__global__ void __launch_bounds__(256) trace_planes_kernel(const float4* __restrict__ spheres,
                                                           const int32_t* __restrict__ adjacency,
                                                           const int32_t* __restrict__ adj_off, int n_sites,
                                                           int64_t n_faces, float4* __restrict__ planes) {
#pragma clang fp contract(off)
  __shared__ int s_off[4][WAVE + 1];
  __shared__ float4 s_sph[4][WAVE];
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int s0 = (blockIdx.x * 4 + wv) * WAVE;
  if (s0 >= n_sites) return;
  const int ns = min(WAVE, n_sites - s0);
  int* so = s_off[wv];
  float4* sp = s_sph[wv];
  if (lane < ns) {
    sp[lane] = spheres[s0 + lane];
    so[lane] = adj_off[s0 + lane];
  }
  if (lane == 0) so[ns] = adj_off[s0 + ns];  // ns may be 64: one past the last lane
  wave_sync();
  // offsets past the face arrays (an inconsistent CSR) end the pass there
  const int e0 = so[0], e1 = (int)min((int64_t)so[ns], n_faces);
  for (int e = e0 + lane; e < e1; e += WAVE) {
    int l = 0;  // largest l with so[l] <= e: the site of face e
#pragma unroll
    for (int step = WAVE / 2; step >= 1; step >>= 1)
      if (l + step < ns && so[l + step] <= e) l += step;
    const float4 cs = sp[l];
    const f3 c{cs.x, cs.y, cs.z};
    const float r = cs.w;
    const float4 a = spheres[adjacency[e]];
    planes[e] = make_float4(a.x - c.x, a.y - c.y, a.z - c.z,
                            0.5f * ((a.x * a.x + a.y * a.y + a.z * a.z) - a.w * a.w) -
                                0.5f * ((c.x * c.x + c.y * c.y + c.z * c.z) - r * r));
  }
}

// Tiles by descending walk length: a counting sort over ORDER_BUCKETS
// lengths (longer walks share the last bucket) in one workgroup.
constexpr int ORDER_BUCKETS = 1024;

__global__ void __launch_bounds__(ORDER_BUCKETS) trace_tile_order_kernel(const int32_t* __restrict__ steps, int n_tiles,
                                                                         int32_t* __restrict__ order) {
  __shared__ int s_cnt[ORDER_BUCKETS];
  __shared__ int s_scan[ORDER_BUCKETS];
  const int t = threadIdx.x;
  s_cnt[t] = 0;
  __syncthreads();
  // bucket 0 holds the longest walks
  auto bucket = [&](int k) { return ORDER_BUCKETS - 1 - min(max(steps[k], 0), ORDER_BUCKETS - 1); };
  for (int k = t; k < n_tiles; k += ORDER_BUCKETS) atomicAdd(&s_cnt[bucket(k)], 1);
  __syncthreads();
  // inclusive scan of the counts (Hillis-Steele over the workgroup)
  int v = s_cnt[t];
  s_scan[t] = v;
  __syncthreads();
  for (int off = 1; off < ORDER_BUCKETS; off <<= 1) {
    const int add = t >= off ? s_scan[t - off] : 0;
    __syncthreads();
    s_scan[t] += add;
    __syncthreads();
  }
  s_cnt[t] = s_scan[t] - v;  // first slot of bucket t
  __syncthreads();
  for (int k = t; k < n_tiles; k += ORDER_BUCKETS) order[atomicAdd(&s_cnt[bucket(k)], 1)] = k;
}

static int check_trace(const pf_trace_args* a, const char* who) {
  if (!a) return set_error(PF_ERR_ARG, "%s: null args", who);
  const pf_scene& s = a->scene;
  if (s.n_sites < 1 || !s.spheres || !s.nsigmas || !s.texel_sites || !s.texel_rgbh || !s.adjacency_offsets ||
      !s.adjacency)
    return set_error(PF_ERR_ARG, "%s: missing scene arrays", who);
  if (s.n_texels < 1 || s.n_texels > PF_MAX_TEXELS) return set_error(PF_ERR_ARG, "%s: n_texels out of range", who);
  if (a->cam.width < 2 || a->cam.height < 2) return set_error(PF_ERR_ARG, "%s: image must be >= 2x2", who);
  if (!a->start_sites && (a->start_site < 0 || a->start_site >= s.n_sites))
    return set_error(PF_ERR_ARG, "%s: start_site %d outside 0..%d", who, a->start_site, s.n_sites - 1);
  if (a->cells && a->max_cells < 1) return set_error(PF_ERR_ARG, "%s: cells needs max_cells >= 1", who);
  return PF_OK;
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_trace_fwd(const pf_trace_args* a, void* stream) {
  int st = check_trace(a, "pf_trace_fwd");
  if (st) return st;
  if (!a->color) return set_error(PF_ERR_ARG, "pf_trace_fwd: missing colour output");
  TraceParams p = make_params(a);
  p.color = a->color;
  p.log_t = a->log_t;
  p.normal = a->normal;
  p.depth = a->depth;
  p.q_step = a->q_step;
  p.q_sigma = a->q_sigma;
  p.cells = a->cells;
  p.K = a->max_cells;
  p.n_cells = a->n_cells;
  p.tile_steps = a->tile_steps;
  if (a->tile_steps && a->compact > 0)
    return set_error(PF_ERR_ARG, "pf_trace_fwd: tile_steps needs one wave per tile (compact = 0)");
  const int fm = p.adj_diff ? FM_HALF : (p.planes ? FM_PLANES : FM_SPHERES);
  hipStream_t hs = as_stream(stream);
  if (a->compact > 0) {
    const int group_tiles = a->compact;
    const int groups = (p.n_tiles + group_tiles - 1) / group_tiles;
    const dim3 grid((groups + TWPB - 1) / TWPB);
    if (fm == FM_HALF)
      trace_fwd_kernel<true, FM_HALF><<<grid, TWAVE * TWPB, 0, hs>>>(p, group_tiles);
    else if (fm == FM_PLANES)
      trace_fwd_kernel<true, FM_PLANES><<<grid, TWAVE * TWPB, 0, hs>>>(p, group_tiles);
    else
      trace_fwd_kernel<true, FM_SPHERES><<<grid, TWAVE * TWPB, 0, hs>>>(p, group_tiles);
  } else {
    const dim3 grid((p.n_tiles + TWPB - 1) / TWPB);
    if (fm == FM_HALF)
      trace_fwd_kernel<false, FM_HALF><<<grid, TWAVE * TWPB, 0, hs>>>(p, 1);
    else if (fm == FM_PLANES)
      trace_fwd_kernel<false, FM_PLANES><<<grid, TWAVE * TWPB, 0, hs>>>(p, 1);
    else
      trace_fwd_kernel<false, FM_SPHERES><<<grid, TWAVE * TWPB, 0, hs>>>(p, 1);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_trace_planes(const pf_scene* scene, int64_t n_faces, float* face_planes, void* stream) {
  if (!scene) return set_error(PF_ERR_ARG, "pf_trace_planes: null scene");
  if (scene->n_sites < 0 || n_faces < 0) return set_error(PF_ERR_ARG, "pf_trace_planes: negative size");
  if (scene->n_sites == 0 || n_faces == 0) return PF_OK;
  if (!scene->spheres || !scene->adjacency_offsets || !scene->adjacency || !face_planes)
    return set_error(PF_ERR_ARG, "pf_trace_planes: needs spheres, the adjacency and the output");
  trace_planes_kernel<<<(scene->n_sites + 4 * WAVE - 1) / (4 * WAVE), 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const float4*>(scene->spheres), scene->adjacency, scene->adjacency_offsets, scene->n_sites,
      n_faces, reinterpret_cast<float4*>(face_planes));
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_trace_tile_order(const int32_t* tile_steps, int32_t n_tiles, int32_t* tile_order, void* stream) {
  if (n_tiles < 0) return set_error(PF_ERR_ARG, "pf_trace_tile_order: n_tiles < 0");
  if (n_tiles == 0) return PF_OK;
  if (!tile_steps || !tile_order) return set_error(PF_ERR_ARG, "pf_trace_tile_order: a NULL array");
  trace_tile_order_kernel<<<1, ORDER_BUCKETS, 0, as_stream(stream)>>>(tile_steps, n_tiles, tile_order);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_trace_check_f64(const pf_trace_check_args* c, void* stream) {
  if (!c) return set_error(PF_ERR_ARG, "pf_trace_check_f64: null args");
  const pf_scene& s = c->scene;
  if (s.n_sites < 1 || !s.spheres || !s.adjacency_offsets || !s.adjacency)
    return set_error(PF_ERR_ARG, "pf_trace_check_f64: needs spheres and the full adjacency");
  if (c->cam.width < 2 || c->cam.height < 2) return set_error(PF_ERR_ARG, "pf_trace_check_f64: image must be >= 2x2");
  if (c->n_rays < 0 || c->max_cells < 1) return set_error(PF_ERR_ARG, "pf_trace_check_f64: bad sizes");
  if (c->n_rays == 0) return PF_OK;
  if (!c->pixels || !c->cells || !c->n_cells || !c->t_enter || !c->t_exit || !c->next_cell || !c->exit_gap ||
      !c->exit_tol)
    return set_error(PF_ERR_ARG, "pf_trace_check_f64: a NULL array");
  pf_trace_args a{};
  a.cam = c->cam;
  a.scene = s;
  TraceParams p = make_params(&a);
  trace_check_f64_kernel<<<(c->n_rays + 255) / 256, 256, 0, as_stream(stream)>>>(
      p, c->n_rays, c->pixels, c->cells, c->n_cells, c->max_cells, c->t_enter, c->t_exit, c->next_cell, c->exit_gap,
      c->exit_tol);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_trace_bwd(const pf_trace_bwd_args* b, void* stream) {
  if (!b) return set_error(PF_ERR_ARG, "pf_trace_bwd: null args");
  int st = check_trace(&b->fwd, "pf_trace_bwd");
  if (st) return st;
  if (!b->color || !b->normal || !b->log_t || !b->q_step || !b->q_sigma || !b->grad_color || !b->grad_normal ||
      !b->grad_opacity)
    return set_error(PF_ERR_ARG, "pf_trace_bwd: missing saved outputs or incoming gradients");
  if (!b->grad_spheres || !b->grad_nsigmas || !b->grad_texel_sites || !b->grad_texel_rgb || !b->grad_texel_height)
    return set_error(PF_ERR_ARG, "pf_trace_bwd: missing gradient outputs");
  TraceParams p = make_params(&b->fwd);
  p.s_color = b->color;
  p.s_normal = b->normal;
  p.s_log_t = b->log_t;
  p.s_q_step = b->q_step;
  p.s_q_sigma = b->q_sigma;
  p.g_color = b->grad_color;
  p.g_normal = b->grad_normal;
  p.g_depth = b->grad_depth;
  p.g_opacity = b->grad_opacity;
  p.g_spheres = b->grad_spheres;
  p.g_nsigmas = b->grad_nsigmas;
  p.g_tsites = b->grad_texel_sites;
  p.g_trgb = b->grad_texel_rgb;
  p.g_theight = b->grad_texel_height;
  p.tile_order = b->tile_order;
  const dim3 grid((p.n_tiles + TWPB - 1) / TWPB);
  hipStream_t hs = as_stream(stream);
  if (p.adj_diff)
    trace_bwd_kernel<FM_HALF><<<grid, TWAVE * TWPB, 0, hs>>>(p);
  else if (p.planes)
    trace_bwd_kernel<FM_PLANES><<<grid, TWAVE * TWPB, 0, hs>>>(p);
  else
    trace_bwd_kernel<FM_SPHERES><<<grid, TWAVE * TWPB, 0, hs>>>(p);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
