// Cech adjacency on the GPU (the training adjacency of AABBTree.update /
// build_cech_complex, bvh.py:76-123, with the predicate of count_adjacent /
// write_adjacent, bvh.py:7-69): each site's box is lo = c - r, hi = c + r in
// float32; j != i is a neighbour of i when the boxes overlap (the BVH query,
// touching counts) and |c_i - c_j| < r_i + r_j, with centre and radius taken
// back from the box, c = 0.5 (lo + hi), r = 0.5 (hi.x - lo.x).
//
// Instead of a BVH, a hashed grid per radius level. Level l holds the sites
// with r in (R / 2^(l+1), R / 2^l] (R the largest radius; the last level also
// takes every smaller radius) in cells of edge R / 2^l, so a site that looks
// for level-l neighbours within r_i + R / 2^l scans at most 6 cells per axis
// (measured on 1M- and 3M-site scenes: 10% faster than cells twice as wide,
// 30% faster than half as wide). The sites are counting-sorted by hash bucket of (level, cell), and one
// thread per bucket-ordered site (a wave shares its cells and reads them
// coalesced) queries the levels at or above its own radius: a pair of unequal
// levels is found once, by its smaller site, which adds each site to the
// other's list; a pair of one level is found from both sides. The count pass
// writes degrees, the emit pass writes neighbour ids as records into each
// site's segment through a per-site cursor, and the segmented tile-list sort
// then orders every segment ascending (the reference's order is its BVH's
// traversal order, which Warp does not specify). This file is compiled
// without FMA contraction and sqrtf is correctly rounded, so the predicate is
// the float32 one written above, operation for operation.
#include "pf_common.hpp"

namespace pf {

constexpr int CECH_LEVELS = 8;
// Cell edge of level l: R 2^(CECH_EDGE_LOG2 - l).
constexpr int CECH_EDGE_LOG2 = 0;
constexpr int CECH_COORD_BITS = 20;
constexpr int CECH_COORD_MAX = (1 << CECH_COORD_BITS) - 1;
constexpr int CECH_SCAN_TILE = 4096;  // entries per workgroup of the bucket scan
// One wave per workgroup for the pair search: a site's search ranges from a
// few to thousands of candidates, and a workgroup keeps its slots until its
// slowest wave ends.
constexpr int CECH_PAIRS_THREADS = 64;

struct CechGrid {
  float lo[3];
  float rmax;                    // R; zero when no pair can exist
  float edge[CECH_LEVELS];       // cell edge per level
  float inv[CECH_LEVELS];        // 1 / edge
  uint32_t mask;                 // buckets - 1
  int any;                       // 0: no site can have a neighbour
};

struct CechBall {
  float lx, ly, lz, hx, hy, hz;  // the float32 box
  float cx, cy, cz, r;           // centre and radius taken back from it
};

__device__ __forceinline__ CechBall cech_ball(float4 s) {
  CechBall b;
  b.lx = s.x - s.w; b.ly = s.y - s.w; b.lz = s.z - s.w;
  b.hx = s.x + s.w; b.hy = s.y + s.w; b.hz = s.z + s.w;
  b.cx = 0.5f * (b.lx + b.hx); b.cy = 0.5f * (b.ly + b.hy); b.cz = 0.5f * (b.lz + b.hz);
  b.r = 0.5f * (b.hx - b.lx);
  return b;
}

__device__ __forceinline__ bool cech_pair(const CechBall& a, const CechBall& b) {
  if (a.lx > b.hx || b.lx > a.hx || a.ly > b.hy || b.ly > a.hy || a.lz > b.hz || b.lz > a.hz) return false;
  const float dx = a.cx - b.cx, dy = a.cy - b.cy, dz = a.cz - b.cz;
  return sqrtf(dx * dx + dy * dy + dz * dz) < a.r + b.r;
}

// A site whose centre or radius (after the box round trip) is not finite
// fails the predicate against every site: it is left out of the bounds and
// queries nothing.
__device__ __forceinline__ bool cech_finite(const CechBall& b) {
  return isfinite(b.cx) && isfinite(b.cy) && isfinite(b.cz) && isfinite(b.r);
}

// Level of a radius: the largest l < CECH_LEVELS with r * 2^l <= R (scaling
// by powers of two is exact, so r <= R / 2^l holds exactly).
__device__ __forceinline__ int cech_level(float r, float rmax) {
  int l = 0;
  while (l < CECH_LEVELS - 1 && ldexpf(r, l + 1) <= rmax) l++;
  return l;
}

__device__ __forceinline__ int cech_coord(float v, float lo, float inv) {
  const float f = (v - lo) * inv;
  // NaN and negative land in cell 0, beyond the range in the last cell
  return f >= 0.0f ? (f < (float)CECH_COORD_MAX ? (int)f : CECH_COORD_MAX) : 0;
}

// Spreads the low 20 bits of v to every third bit.
__device__ __forceinline__ uint64_t cech_spread(uint32_t v) {
  uint64_t x = v & 0xfffffu;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// Cell key: level in bits 60-62, the Morton code of the cell in bits 0-59.
__device__ __forceinline__ uint64_t cech_key(int l, int x, int y, int z) {
  return ((uint64_t)l << 60) | cech_spread(x) | (cech_spread(y) << 1) | (cech_spread(z) << 2);
}

// Bucket of a cell: 4x4x4-cell Morton blocks stay contiguous (the low six
// bits), the blocks themselves are scattered by the splitmix64 finaliser. In
// bucket order a wave then holds sites of a few neighbouring cells, whose
// queries meet the same buckets.
__device__ __forceinline__ uint32_t cech_hash(uint64_t key, uint32_t mask) {
  uint64_t k = key >> 6;
  k ^= k >> 30;
  k *= 0xbf58476d1ce4e5b9ull;
  k ^= k >> 27;
  k *= 0x94d049bb133111ebull;
  k ^= k >> 31;
  return (uint32_t)((k << 6) | (key & 63)) & mask;
}

// Order-preserving int image of a float, for min/max through integer atomics.
__device__ __forceinline__ int ord_of(float f) {
  const int i = __builtin_bit_cast(int, f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float float_of(int i) { return __builtin_bit_cast(float, i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void __launch_bounds__(256) cech_fill_kernel(int* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void cech_init_kernel(int* __restrict__ acc) {
  const int t = threadIdx.x;
  if (t < 3) acc[t] = 0x7fffffff;
  else if (t < 6) acc[t] = (int)0x80000000;
  else if (t < 8 + CECH_LEVELS) acc[t] = 0;  // rmax, spare, level counts
}

// Bounds of the centres and the largest radius (acc: min xyz, max xyz, rmax).
__global__ void __launch_bounds__(256) cech_bounds_kernel(const float4* __restrict__ spheres, int n,
                                                         int* __restrict__ acc) {
  float mn[3] = {3.0e38f, 3.0e38f, 3.0e38f}, mx[3] = {-3.0e38f, -3.0e38f, -3.0e38f}, rm = 0.0f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const CechBall b = cech_ball(spheres[i]);
    if (!cech_finite(b)) continue;
    mn[0] = fminf(mn[0], b.cx); mn[1] = fminf(mn[1], b.cy); mn[2] = fminf(mn[2], b.cz);
    mx[0] = fmaxf(mx[0], b.cx); mx[1] = fmaxf(mx[1], b.cy); mx[2] = fmaxf(mx[2], b.cz);
    rm = fmaxf(rm, b.r);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], m, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], m, 64));
    }
    rm = fmaxf(rm, __shfl_xor(rm, m, 64));
  }
  // one atomic per value per workgroup: the waves meet in LDS first
  __shared__ float s_b[4][7];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    for (int k = 0; k < 3; k++) {
      s_b[w][k] = mn[k];
      s_b[w][3 + k] = mx[k];
    }
    s_b[w][6] = rm;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    atomicMin(&acc[k], ord_of(fminf(fminf(s_b[0][k], s_b[1][k]), fminf(s_b[2][k], s_b[3][k]))));
  } else if (threadIdx.x < 7) {
    const int k = threadIdx.x;
    atomicMax(&acc[k], ord_of(fmaxf(fmaxf(s_b[0][k], s_b[1][k]), fmaxf(s_b[2][k], s_b[3][k]))));
  }
}

// Cell edges per level: R / 2^l, raised where needed so that the span of
// the centres fits CECH_COORD_MAX cells and an edge is at least 1e-5 of the
// largest coordinate (the query margin below is 3e-6 of it, so a query spans
// at most 6 cells per axis). No finite site, or R = 0, means no pair.
__global__ void cech_grid_kernel(const int* __restrict__ acc, CechGrid* __restrict__ g, uint32_t mask) {
  if (threadIdx.x != 0) return;
  float lo[3], span = 0.0f, cmax = 0.0f;
  bool finite = true;
  for (int k = 0; k < 3; k++) {
    lo[k] = float_of(acc[k]);
    const float hi = float_of(acc[3 + k]);
    finite = finite && lo[k] <= hi;  // false when no site was finite
    span = fmaxf(span, hi - lo[k]);
    cmax = fmaxf(cmax, fmaxf(fabsf(lo[k]), fabsf(hi)));
  }
  const float rmax = float_of(acc[6]);
  const bool any = finite && rmax > 0.0f;
  const float min_edge = fmaxf(span / (float)(CECH_COORD_MAX - 1), 1e-5f * cmax);
  for (int k = 0; k < 3; k++) g->lo[k] = finite ? lo[k] : 0.0f;
  for (int l = 0; l < CECH_LEVELS; l++) {
    const float e = fmaxf(ldexpf(rmax, CECH_EDGE_LOG2 - l), min_edge);
    g->edge[l] = e;
    g->inv[l] = (any && e > 0.0f) ? 1.0f / e : 0.0f;
  }
  g->rmax = any ? rmax : 0.0f;
  g->mask = mask;
  g->any = any ? 1 : 0;
}

__global__ void __launch_bounds__(256) cech_hist_kernel(const float4* __restrict__ spheres, int n,
                                                       const CechGrid* __restrict__ gp, uint32_t* __restrict__ site_bucket,
                                                       int* __restrict__ bucket_start, int* __restrict__ level_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CechGrid& g = *gp;
  const CechBall b = cech_ball(spheres[i]);
  const int l = cech_level(b.r, g.rmax);
  const float inv = g.inv[l];
  const uint64_t key = cech_key(l, cech_coord(b.cx, g.lo[0], inv), cech_coord(b.cy, g.lo[1], inv),
                                cech_coord(b.cz, g.lo[2], inv));
  const uint32_t h = cech_hash(key, g.mask);
  site_bucket[i] = h;
  atomicAdd(&bucket_start[h + 1], 1);
  // one atomic per level present in the wave instead of one per site
  uint64_t left = __ballot(true);
  while (left) {
    const int lead = __builtin_ctzll(left);
    const int ll = __builtin_amdgcn_readlane(l, lead);
    const uint64_t same = __ballot(l == ll) & left;
    if ((threadIdx.x & 63) == lead) atomicAdd(&level_count[ll], (int)__builtin_popcountll(same));
    left &= ~same;
  }
}

// Inclusive scan of a[0..n) in place, three phases: per-tile scans with the
// tile totals in part[], one workgroup scans part[] (exclusive), then every
// tile adds its prefix.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Exclusive prefix of v over the workgroup (blockDim.x = 1024) and the total.
__device__ __forceinline__ int block_excl_prefix(int v, int* s_wave, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  if (w == 0) {
    const int t = lane < 16 ? s_wave[lane] : 0;
    const int ti = wave_incl_scan(t);
    if (lane < 16) s_wave[16 + lane] = ti - t;
    if (lane == 15) s_wave[32] = ti;
  }
  __syncthreads();
  total = s_wave[32];
  const int r = s_wave[16 + w] + inc - v;
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(1024) cech_scan_tiles_kernel(int* __restrict__ a, int64_t n, int* __restrict__ part) {
  __shared__ int s_wave[33];
  const int64_t base = (int64_t)blockIdx.x * CECH_SCAN_TILE + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    v[m] = base + m < n ? a[base + m] : 0;
    s += v[m];
  }
  int total;
  int run = block_excl_prefix(s, s_wave, total);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    run += v[m];
    if (base + m < n) a[base + m] = run;
  }
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) cech_scan_parts_kernel(int* __restrict__ part, int np) {
  __shared__ int s_wave[33];
  const int per = (np + 1023) / 1024;
  const int a = min((int)threadIdx.x * per, np), b = min(a + per, np);
  int s = 0;
  for (int i = a; i < b; i++) s += part[i];
  int total;
  int run = block_excl_prefix(s, s_wave, total);
  for (int i = a; i < b; i++) {
    const int v = part[i];
    part[i] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(1024) cech_scan_add_kernel(int* __restrict__ a, int64_t n, const int* __restrict__ part) {
  const int add = part[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * CECH_SCAN_TILE + threadIdx.x * 4;
#pragma unroll
  for (int m = 0; m < 4; m++)
    if (base + m < n) a[base + m] += add;
}

// Balls, ids and cell keys in bucket order (order inside a bucket is atomic
// order; the emitted lists are sorted afterwards, so nothing depends on it).
__global__ void __launch_bounds__(256) cech_scatter_kernel(const float4* __restrict__ spheres, int n,
                                                          const CechGrid* __restrict__ gp,
                                                          const uint32_t* __restrict__ site_bucket,
                                                          const int* __restrict__ bucket_start,
                                                          int* __restrict__ bucket_fill, float4* __restrict__ ord_balls,
                                                          int* __restrict__ ord_ids, uint64_t* __restrict__ ord_keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CechGrid& g = *gp;
  const uint32_t h = site_bucket[i];
  const int k = bucket_start[h] + atomicAdd(&bucket_fill[h], 1);
  const float4 s = spheres[i];
  const CechBall b = cech_ball(s);
  const int l = cech_level(b.r, g.rmax);
  const float inv = g.inv[l];
  ord_balls[k] = s;
  ord_ids[k] = i;
  ord_keys[k] = cech_key(l, cech_coord(b.cx, g.lo[0], inv), cech_coord(b.cy, g.lo[1], inv),
                         cech_coord(b.cz, g.lo[2], inv));
}

// Pairs of one query thread. A site's segment holds first the neighbours its
// own thread finds (same or coarser level; written in order, no atomics),
// then those that finer sites add (through a per-site atomic cursor). COUNT
// writes the degree (atomics for the added part) and the own part's length;
// otherwise the ids go out as records, and entries beyond a segment are
// counted into *dropped, never written.
template <bool COUNT>
__global__ void __launch_bounds__(CECH_PAIRS_THREADS)
    cech_pairs_kernel(int n, const CechGrid* __restrict__ gp, const int* __restrict__ level_count,
                      const int* __restrict__ bucket_start, const float4* __restrict__ ord_balls,
                      const int* __restrict__ ord_ids, const uint64_t* __restrict__ ord_keys,
                      int32_t* __restrict__ counts, int32_t* __restrict__ own_deg, const int64_t* __restrict__ offsets,
                      int32_t* __restrict__ fill, uint64_t* __restrict__ records, int32_t* __restrict__ dropped) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !gp->any) return;
  const CechGrid& g = *gp;
  const int i = ord_ids[k];
  const CechBall bi = cech_ball(ord_balls[k]);
  if (!cech_finite(bi)) {
    if (COUNT) own_deg[i] = 0;
    return;
  }
  const int li = (int)(ord_keys[k] >> 60);
  int own = 0, lost = 0;
  int64_t w = 0, w_end = 0;
  if (!COUNT) {
    w = offsets[i];
    w_end = min(offsets[i + 1], w + own_deg[i]);
  }
  for (int l = 0; l <= li; l++) {
    if (level_count[l] == 0) continue;
    const float inv = g.inv[l];
    // every level-l neighbour centre lies within r_i + R / 2^l; the margin
    // covers the rounding of the bound and of the cell coordinates
    const float reach0 = bi.r + ldexpf(g.rmax, -l);
    const float reach = reach0 * 1.001f + 1e-6f * (fabsf(bi.cx) + fabsf(bi.cy) + fabsf(bi.cz));
    const int x0 = cech_coord(bi.cx - reach, g.lo[0], inv), x1 = cech_coord(bi.cx + reach, g.lo[0], inv);
    const int y0 = cech_coord(bi.cy - reach, g.lo[1], inv), y1 = cech_coord(bi.cy + reach, g.lo[1], inv);
    const int z0 = cech_coord(bi.cz - reach, g.lo[2], inv), z1 = cech_coord(bi.cz + reach, g.lo[2], inv);
    for (int z = z0; z <= z1; z++)
      for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) {
          const uint64_t key = cech_key(l, x, y, z);
          const uint32_t h = cech_hash(key, g.mask);
          const int q1 = bucket_start[h + 1];
          for (int q = bucket_start[h]; q < q1; q++) {
            if (ord_keys[q] != key || q == k) continue;
            if (!cech_pair(bi, cech_ball(ord_balls[q]))) continue;
            const int j = ord_ids[q];
            if (COUNT) {
              own++;
              if (l < li) atomicAdd(&counts[j + 1], 1);
            } else {
              if (w < w_end) records[w++] = (uint64_t)(uint32_t)j;
              else lost++;
              if (l < li) {
                const int64_t s = offsets[j] + own_deg[j] + atomicAdd(&fill[j], 1);
                if (s < offsets[j + 1]) records[s] = (uint64_t)(uint32_t)i;
                else lost++;
              }
            }
          }
        }
  }
  if (COUNT) {
    own_deg[i] = own;
    if (own) atomicAdd(&counts[i + 1], own);
  } else if (lost) {
    atomicAdd(dropped, lost);
  }
}

struct CechWorkspace {
  CechGrid* grid;
  int* acc;            // bounds words, then CECH_LEVELS level counts
  int* bucket_start;   // [buckets + 1]
  int* bucket_fill;    // [buckets]
  int* part;           // [scan tiles]
  uint32_t* site_bucket;  // [n]
  int* ord_ids;        // [n]
  int* fill;           // [n]
  int* own_deg;        // [n]
  uint64_t* ord_keys;  // [n]
  float4* ord_balls;   // [n]
  int64_t buckets;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static int64_t cech_buckets(int64_t n) {
  int64_t b = 1024;
  while (b < 2 * n) b <<= 1;
  return b;
}

static int64_t cech_scan_tiles(int64_t buckets) { return (buckets + 1 + CECH_SCAN_TILE - 1) / CECH_SCAN_TILE; }

static size_t cech_ws_bytes(int64_t n) {
  const int64_t nb = cech_buckets(n);
  return align256(sizeof(CechGrid)) + align256(sizeof(int) * (8 + CECH_LEVELS)) +
         align256(sizeof(int) * (size_t)(nb + 1)) + align256(sizeof(int) * (size_t)nb) +
         align256(sizeof(int) * (size_t)cech_scan_tiles(nb)) + 4 * align256(sizeof(int) * (size_t)n) +
         align256(sizeof(uint64_t) * (size_t)n) + align256(sizeof(float4) * (size_t)n);
}

static CechWorkspace cech_ws(void* base, int64_t n) {
  char* p = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += align256(bytes);
    return q;
  };
  CechWorkspace w;
  w.buckets = cech_buckets(n);
  w.grid = reinterpret_cast<CechGrid*>(take(sizeof(CechGrid)));
  w.acc = reinterpret_cast<int*>(take(sizeof(int) * (8 + CECH_LEVELS)));
  w.bucket_start = reinterpret_cast<int*>(take(sizeof(int) * (size_t)(w.buckets + 1)));
  w.bucket_fill = reinterpret_cast<int*>(take(sizeof(int) * (size_t)w.buckets));
  w.part = reinterpret_cast<int*>(take(sizeof(int) * (size_t)cech_scan_tiles(w.buckets)));
  w.site_bucket = reinterpret_cast<uint32_t*>(take(sizeof(int) * (size_t)n));
  w.ord_ids = reinterpret_cast<int*>(take(sizeof(int) * (size_t)n));
  w.fill = reinterpret_cast<int*>(take(sizeof(int) * (size_t)n));
  w.own_deg = reinterpret_cast<int*>(take(sizeof(int) * (size_t)n));
  w.ord_keys = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * (size_t)n));
  w.ord_balls = reinterpret_cast<float4*>(take(sizeof(float4) * (size_t)n));
  return w;
}

static int cech_check(const char* fn, const float* spheres, int32_t n_sites, const void* workspace,
                      size_t workspace_bytes) {
  if (n_sites < 0 || n_sites > PF_CECH_MAX_SITES || (n_sites > 0 && (!spheres || !workspace)))
    return set_error(PF_ERR_ARG, "%s: bad arguments", fn);
  if (reinterpret_cast<uintptr_t>(spheres) % 16 || reinterpret_cast<uintptr_t>(workspace) % 256)
    return set_error(PF_ERR_ARG, "%s: spheres must be 16-byte and workspace 256-byte aligned", fn);
  if (workspace_bytes < cech_ws_bytes(n_sites))
    return set_error(PF_ERR_OVERFLOW, "%s: workspace of %zu bytes below the %zu that pf_cech_workspace_size gives", fn,
                     workspace_bytes, cech_ws_bytes(n_sites));
  return PF_OK;
}

static unsigned cech_blocks(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace pf

using namespace pf;

extern "C" {

size_t pf_cech_workspace_size(int32_t n_sites) {
  if (n_sites < 0 || n_sites > PF_CECH_MAX_SITES) return 0;
  return cech_ws_bytes(n_sites);
}

int pf_cech_count(const float* spheres, int32_t n_sites, void* workspace, size_t workspace_bytes, int32_t* counts,
                  void* stream) {
  if (!counts) return set_error(PF_ERR_ARG, "pf_cech_count: bad arguments");
  const int s = cech_check("pf_cech_count", spheres, n_sites, workspace, workspace_bytes);
  if (s != PF_OK) return s;
  hipStream_t st = as_stream(stream);
  cech_fill_kernel<<<cech_blocks((int64_t)n_sites + 1, 256), 256, 0, st>>>(counts, (int64_t)n_sites + 1);
  if (n_sites > 0) {
    const float4* sp = reinterpret_cast<const float4*>(spheres);
    const CechWorkspace w = cech_ws(workspace, n_sites);
    const unsigned blocks = cech_blocks(n_sites, 256);
    const int64_t nb1 = w.buckets + 1;
    const unsigned tiles = (unsigned)cech_scan_tiles(w.buckets);
    cech_init_kernel<<<1, 64, 0, st>>>(w.acc);
    cech_bounds_kernel<<<min(blocks, 512u), 256, 0, st>>>(sp, n_sites, w.acc);
    cech_grid_kernel<<<1, 64, 0, st>>>(w.acc, w.grid, (uint32_t)(w.buckets - 1));
    cech_fill_kernel<<<cech_blocks(nb1, 256), 256, 0, st>>>(w.bucket_start, nb1);
    cech_fill_kernel<<<cech_blocks(w.buckets, 256), 256, 0, st>>>(w.bucket_fill, w.buckets);
    cech_hist_kernel<<<blocks, 256, 0, st>>>(sp, n_sites, w.grid, w.site_bucket, w.bucket_start, w.acc + 8);
    cech_scan_tiles_kernel<<<tiles, 1024, 0, st>>>(w.bucket_start, nb1, w.part);
    cech_scan_parts_kernel<<<1, 1024, 0, st>>>(w.part, (int)tiles);
    cech_scan_add_kernel<<<tiles, 1024, 0, st>>>(w.bucket_start, nb1, w.part);
    cech_scatter_kernel<<<blocks, 256, 0, st>>>(sp, n_sites, w.grid, w.site_bucket, w.bucket_start, w.bucket_fill,
                                                w.ord_balls, w.ord_ids, w.ord_keys);
    cech_pairs_kernel<true><<<cech_blocks(n_sites, CECH_PAIRS_THREADS), CECH_PAIRS_THREADS, 0, st>>>(n_sites, w.grid, w.acc + 8, w.bucket_start, w.ord_balls,
                                                    w.ord_ids, w.ord_keys, counts, w.own_deg, nullptr, nullptr,
                                                    nullptr, nullptr);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_cech_emit(const float* spheres, int32_t n_sites, void* workspace, size_t workspace_bytes,
                 const int64_t* offsets, uint64_t* records, int32_t* dropped, void* stream) {
  if (n_sites > 0 && (!offsets || !dropped)) return set_error(PF_ERR_ARG, "pf_cech_emit: bad arguments");
  const int s = cech_check("pf_cech_emit", spheres, n_sites, workspace, workspace_bytes);
  if (s != PF_OK) return s;
  if (n_sites == 0) return PF_OK;
  hipStream_t st = as_stream(stream);
  const CechWorkspace w = cech_ws(workspace, n_sites);
  const unsigned blocks = cech_blocks(n_sites, 256);
  cech_fill_kernel<<<blocks, 256, 0, st>>>(w.fill, n_sites);
  cech_pairs_kernel<false><<<cech_blocks(n_sites, CECH_PAIRS_THREADS), CECH_PAIRS_THREADS, 0, st>>>(n_sites, w.grid, w.acc + 8, w.bucket_start, w.ord_balls, w.ord_ids,
                                                   w.ord_keys, nullptr, w.own_deg, offsets, w.fill, records, dropped);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
