// Restricted power adjacency on the GPU: the Cech pairs whose power face
// meets the balls (no reference kernel; the reference builds its power
// adjacency on the CPU from the lifted convex hull, benchmark.py:25-83, and
// filters it by |c_i - c_j| < r_i + r_j, which keeps faces that lie outside
// both balls and that no ray inside a ball can reach).
//
// On the radical plane of a and b the two balls cut one disk; the face meets
// the balls when the disk keeps a point with pow_a <= pow_k for every other
// site k, and only a's Cech neighbours can cut the disk. Per canonical edge
// (a < b), in FP64: the plane frame (u along c_b - c_a, in-plane e1, e2, the
// plane offset s, the disk radius^2 rho2), each neighbour k a half-plane
// a1 y1 + a2 y2 <= beta (only sites overlapping both balls can cut the disk),
// and Seidel's incremental LP (smallest y1; a violated constraint moves the
// optimum onto its line, a chord of the disk clipped by the earlier
// constraints). Both directed entries take the answer. A site left with no
// face lies wholly inside or wholly outside its cell; when its centre is
// power-nearer to some neighbour than to itself (w < 0 below) it is outside,
// and it keeps, one way, every neighbour whose radical plane cuts into its
// ball, so it renders empty.
//
// One workgroup of four waves per site, threads over its edges to higher ids,
// with a's (c_k - c_a, w_k, r_k, id) staged in LDS up to POWER_LDS_CAP sites;
// the constraints that cannot reach the disk (balls missing b) cost one test
// each. The work per site grows with the square of its degree, so the few
// sites with thousands of neighbours set the time. Measured on the 1M-site
// GMM scene (134M Cech entries): one wave per site 790 ms, four 340-380 ms;
// reading the shorter of the two lists from global memory instead was
// slower (410-430 ms). Compiled without FMA contraction, with correctly
// rounded FP64 division and square root, in the operation order of the CPU
// restatement, so both give the same bits.
#include "pf_common.hpp"

namespace pf {

constexpr int POWER_THREADS = 256;
constexpr int POWER_LDS_CAP = 1024;

struct PwEdge {
  double u[3], e1[3], e2[3], s, rho2;
};

__device__ __forceinline__ double pw_dot(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

__device__ __forceinline__ bool pw_frame(float4 sa, float4 sb, PwEdge& f) {
  double d[3] = {(double)sb.x - (double)sa.x, (double)sb.y - (double)sa.y, (double)sb.z - (double)sa.z};
  const double L2 = pw_dot(d, d);
  if (!(L2 > 0.0)) return false;
  const double L = sqrt(L2);
  const double ra2 = (double)sa.w * (double)sa.w, rb2 = (double)sb.w * (double)sb.w;
  const double s = ((L2 + ra2) - rb2) / (2.0 * L);
  const double rho2 = ra2 - s * s;
  if (!(rho2 > 0.0)) return false;
  for (int k = 0; k < 3; k++) f.u[k] = d[k] / L;
  const double ax = fabs(f.u[0]), ay = fabs(f.u[1]), az = fabs(f.u[2]);
  double t[3] = {0.0, 0.0, 0.0};
  if (ax <= ay && ax <= az) t[0] = 1.0;
  else if (ay <= az) t[1] = 1.0;
  else t[2] = 1.0;
  double e1[3] = {f.u[1] * t[2] - f.u[2] * t[1], f.u[2] * t[0] - f.u[0] * t[2], f.u[0] * t[1] - f.u[1] * t[0]};
  const double n1 = sqrt(pw_dot(e1, e1));
  for (int k = 0; k < 3; k++) f.e1[k] = e1[k] / n1;
  f.e2[0] = f.u[1] * f.e1[2] - f.u[2] * f.e1[1];
  f.e2[1] = f.u[2] * f.e1[0] - f.u[0] * f.e1[2];
  f.e2[2] = f.u[0] * f.e1[1] - f.u[1] * f.e1[0];
  f.s = s;
  f.rho2 = rho2;
  return true;
}

// d = c_k - c_a, w = (|d|^2 + r_a^2 - r_k^2) / 2
__device__ __forceinline__ double4 pw_site(float4 sa, float4 sk) {
  double d[3] = {(double)sk.x - (double)sa.x, (double)sk.y - (double)sa.y, (double)sk.z - (double)sa.z};
  const double ra2 = (double)sa.w * (double)sa.w, rk2 = (double)sk.w * (double)sk.w;
  return make_double4(d[0], d[1], d[2], ((pw_dot(d, d) + ra2) - rk2) * 0.5);
}

// Neighbour ids outside [0, n) stand for the site itself: a zero half-plane
// that never binds, and no face.
__device__ __forceinline__ float4 pw_fetch(const float4* spheres, int k, int n, int a) {
  return spheres[(unsigned)k < (unsigned)n ? k : a];
}

__device__ __forceinline__ void pw_halfplane(const PwEdge& f, double4 dw, double& a1, double& a2, double& beta) {
  const double d[3] = {dw.x, dw.y, dw.z};
  a1 = pw_dot(f.e1, d);
  a2 = pw_dot(f.e2, d);
  beta = dw.w - f.s * pw_dot(f.u, d);
}

// Can site k (one of a's neighbours) cut the disk of (a, b)? The disk lies in
// ball b, so only a ball that overlaps b can; the margin keeps every such
// ball despite rounding. dw = (c_k - c_a, w), db = c_b - c_a.
__device__ __forceinline__ bool pw_reaches(double4 dw, const double* db, double rk, double rb) {
  const double e[3] = {dw.x - db[0], dw.y - db[1], dw.z - db[2]};
  const double s = rk + rb;
  return pw_dot(e, e) < (s * s) * (1.0 + 1e-12);
}

// a's neighbour list as the LP reads it: the first POWER_LDS_CAP sites staged
// in LDS as (c_k - c_a, w_k), r_k and id, the rest from global memory.
struct PwList {
  const int32_t* ids;
  int64_t m;
  const double4* lds_dw;
  const float* lds_r;
  const int* lds_id;
  __device__ __forceinline__ int id(int64_t q) const { return q < POWER_LDS_CAP ? lds_id[q] : ids[q]; }
  __device__ __forceinline__ void site(int64_t q, int k, const float4* spheres, int n, int a, float4 sa,
                                       double4& dw, double& rk) const {
    if (q < POWER_LDS_CAP) {
      dw = lds_dw[q];
      rk = (double)lds_r[q];
    } else {
      const float4 sk = pw_fetch(spheres, k, n, a);
      dw = pw_site(sa, sk);
      rk = (double)sk.w;
    }
  }
};

// Does the face of (a, b) meet the balls? Constraints from list L in order.
__device__ bool pw_face(const float4* __restrict__ spheres, int n, int a, int b, float4 sa, float4 sb,
                        const PwList& L) {
  PwEdge f;
  if (!pw_frame(sa, sb, f)) return false;
  const double db[3] = {(double)sb.x - (double)sa.x, (double)sb.y - (double)sa.y, (double)sb.z - (double)sa.z};
  const double rb = (double)sb.w;
  double y1 = -sqrt(f.rho2), y2 = 0.0;
  for (int64_t q = 0; q < L.m; q++) {
    const int k = L.id(q);
    if (k == b) continue;
    double4 dw;
    double rk;
    L.site(q, k, spheres, n, a, sa, dw, rk);
    if (!pw_reaches(dw, db, rk, rb)) continue;
    double a1, a2, beta;
    pw_halfplane(f, dw, a1, a2, beta);
    if (a1 * y1 + a2 * y2 <= beta) continue;
    const double nn = a1 * a1 + a2 * a2;
    if (!(nn > 0.0)) return false;
    const double x1 = a1 * beta / nn, x2 = a2 * beta / nn;
    const double h2 = f.rho2 - (x1 * x1 + x2 * x2);
    if (h2 < 0.0) return false;
    const double h = sqrt(h2), inv = 1.0 / sqrt(nn);
    const double t1 = -a2 * inv, t2 = a1 * inv;
    double lo = -h, hi = h;
    for (int64_t p = 0; p < q; p++) {
      const int kk = L.id(p);
      if (kk == b) continue;
      double4 dd;
      double rkk;
      L.site(p, kk, spheres, n, a, sa, dd, rkk);
      if (!pw_reaches(dd, db, rkk, rb)) continue;
      double b1, b2, bb;
      pw_halfplane(f, dd, b1, b2, bb);
      const double den = b1 * t1 + b2 * t2;
      const double num = bb - (b1 * x1 + b2 * x2);
      if (den > 0.0) {
        const double v = num / den;
        if (v < hi) hi = v;
      } else if (den < 0.0) {
        const double v = num / den;
        if (v > lo) lo = v;
      } else if (num < 0.0) {
        return false;
      }
      if (lo > hi) return false;
    }
    const double t = t1 > 0.0 ? lo : hi;
    y1 = x1 + t * t1;
    y2 = x2 + t * t2;
  }
  return true;
}

__global__ void __launch_bounds__(256) power_fill_kernel(uint8_t* __restrict__ keep, int64_t m,
                                                         int32_t* __restrict__ counts, int n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) keep[i] = 0;
  if (i <= n) counts[i] = 0;
}

__global__ void __launch_bounds__(POWER_THREADS) power_faces_kernel(const float4* __restrict__ spheres, int n,
                                                                   const int32_t* __restrict__ adj,
                                                                   const int32_t* __restrict__ off, int64_t n_edges,
                                                                   uint8_t* __restrict__ keep,
                                                                   int32_t* __restrict__ counts) {
  const int a = blockIdx.x;
  const int64_t o0 = min((int64_t)off[a], n_edges), o1 = min((int64_t)off[a + 1], n_edges);
  const int64_t m = o1 - o0;
  if (m <= 0) return;
  __shared__ double4 s_dw[POWER_LDS_CAP];
  __shared__ float s_r[POWER_LDS_CAP];
  __shared__ int s_id[POWER_LDS_CAP];
  const float4 sa = spheres[a];
  for (int64_t q = threadIdx.x; q < min(m, (int64_t)POWER_LDS_CAP); q += POWER_THREADS) {
    const int k = adj[o0 + q];
    const float4 sk = pw_fetch(spheres, k, n, a);
    s_dw[q] = pw_site(sa, sk);
    s_r[q] = sk.w;
    s_id[q] = k;
  }
  __syncthreads();
  const PwList La{adj + o0, m, s_dw, s_r, s_id};
  int own = 0;
  for (int64_t p = threadIdx.x; p < m; p += POWER_THREADS) {
    const int b = adj[o0 + p];
    if (b <= a || b >= n) continue;
    if (!pw_face(spheres, n, a, b, sa, spheres[b], La)) continue;
    const int64_t b0 = min((int64_t)off[b], n_edges), b1 = min((int64_t)off[b + 1], n_edges);
    const int64_t mb = max(b1 - b0, (int64_t)0);
    keep[o0 + p] = 1;
    own++;
    // a in b's ascending list
    int64_t lo = b0, hi = b0 + mb;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (adj[mid] < a) lo = mid + 1;
      else hi = mid;
    }
    if (lo < b0 + mb && adj[lo] == a) {
      keep[lo] = 1;
      atomicAdd(&counts[b + 1], 1);
    }
  }
  if (own) atomicAdd(&counts[a + 1], own);
}

// Sites without a face whose cell misses their ball (after power_faces_kernel).
__global__ void __launch_bounds__(256) power_dead_kernel(const float4* __restrict__ spheres, int n,
                                                        const int32_t* __restrict__ adj,
                                                        const int32_t* __restrict__ off, int64_t n_edges,
                                                        uint8_t* __restrict__ keep, int32_t* __restrict__ counts) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n || counts[a + 1] != 0) return;
  const int64_t o0 = min((int64_t)off[a], n_edges), o1 = min((int64_t)off[a + 1], n_edges);
  const float4 sa = spheres[a];
  bool dead = false;
  for (int64_t p = o0; p < o1 && !dead; p++) dead = pw_site(sa, pw_fetch(spheres, adj[p], n, a)).w < 0.0;
  if (!dead) return;
  const double ra = (double)sa.w;
  int c = 0;
  for (int64_t p = o0; p < o1; p++) {
    const double4 dw = pw_site(sa, pw_fetch(spheres, adj[p], n, a));
    const double d[3] = {dw.x, dw.y, dw.z};
    const double L = sqrt(pw_dot(d, d));
    if (L > 0.0 && dw.w < ra * L) {
      keep[p] = 1;
      c++;
    }
  }
  counts[a + 1] = c;
}

// ------------------------------------------------------- face pruning
// Faces of a site's own list that can never bound a ray's interval inside
// its ball (no reference counterpart; the compositing kernels then test
// fewer faces per walked pair). With c the centre, face k is the plane
// n_k.(x - c) = w_k, w_k = ((|n_k|^2 + r^2) - r_k^2)/2, the cell side
// n_k.(x - c) < w_k. Face f is dropped when c is strictly on the cell side of
// f (w_f > 0) and either the plane misses the ball by the margin
// (w_f^2 >= r^2 (1+m)^2 |n_f|^2) or the disc it cuts from the ball lies
// outside another face g with w_g > 0 by m r:
//   n_g.q_f - rho_f |n_g - (n_g.n_f / |n_f|^2) n_f| - w_g > m r |n_g|,
// q_f - c = n_f w_f / |n_f|^2 the disc centre, rho_f its radius. Then the
// ball's part on g's cell side is connected, holds c and misses f's plane,
// so it lies on f's cell side: removing f leaves cell and ball intersected
// exactly as before (and a chain of such removals too, since a witness's
// own witness bounds the same region). FP32 without contraction, in the
// order of the CPU restatement; the margin m covers the compositing's
// rounding. Eight lanes per site (a lane takes every eighth face and tests
// it against all of the site's faces), kept faces first in list order.
constexpr int PRUNE_LANES = 8;

struct PruneFace {
  float nx, ny, nz, w;
};

__device__ __forceinline__ PruneFace prune_face(const int32_t* adjacency, const uint16_t* diff, const float4* spheres,
                                                int64_t e, float4 c) {
  float nx, ny, nz, ra;
  if (diff) {
    const uint2 raw = *reinterpret_cast<const uint2*>(diff + 4 * e);
    nx = half_bits_to_float((uint16_t)(raw.x & 0xffff));
    ny = half_bits_to_float((uint16_t)(raw.x >> 16));
    nz = half_bits_to_float((uint16_t)(raw.y & 0xffff));
    ra = half_bits_to_float((uint16_t)(raw.y >> 16));
  } else {
    const float4 a = spheres[adjacency[e]];
    nx = a.x - c.x;
    ny = a.y - c.y;
    nz = a.z - c.z;
    ra = a.w;
  }
  const float n2 = (nx * nx + ny * ny) + nz * nz;
  return PruneFace{nx, ny, nz, 0.5f * ((n2 + c.w * c.w) - ra * ra)};
}

__global__ void __launch_bounds__(256)
    prune_faces_kernel(const float4* __restrict__ spheres, int n_sites, const int32_t* __restrict__ adjacency,
                       const int32_t* __restrict__ offsets, int64_t n_edges, const uint16_t* __restrict__ diff, float m,
                       int32_t* __restrict__ adj_out, uint16_t* __restrict__ diff_out, int32_t* __restrict__ count) {
  const int lane = threadIdx.x & 63, j = lane & (PRUNE_LANES - 1);
  const int s = (blockIdx.x * blockDim.x + threadIdx.x) / PRUNE_LANES;
  if (s >= n_sites) return;
  // offsets past the edge arrays (an inconsistent CSR) end the range there
  const int a1 = (int)min((int64_t)offsets[s + 1], n_edges), a0 = min(offsets[s], a1);
  const float4 c = spheres[s];
  const float r2 = c.w * c.w, lim = (1.0f + m) * (1.0f + m), mr = m * c.w;
  const int shift = lane & ~(PRUNE_LANES - 1);
  int kept = 0, dropped = 0;
  for (int f0 = a0; f0 < a1; f0 += PRUNE_LANES) {
    const int f = f0 + j;
    bool drop = false;
    if (f < a1) {
      const PruneFace F = prune_face(adjacency, diff, spheres, f, c);
      if (F.w > 0.0f) {
        const float nf2 = (F.nx * F.nx + F.ny * F.ny) + F.nz * F.nz;
        if (F.w * F.w >= (r2 * lim) * nf2) {
          drop = true;
        } else {
          const float q = F.w / nf2;
          const float rho = sqrtf(fmaxf(r2 - F.w * q, 0.0f));
          for (int g = a0; g < a1 && !drop; g++) {
            if (g == f) continue;
            const PruneFace G = prune_face(adjacency, diff, spheres, g, c);
            if (!(G.w > 0.0f)) continue;
            const float ng2 = (G.nx * G.nx + G.ny * G.ny) + G.nz * G.nz;
            const float d = (G.nx * F.nx + G.ny * F.ny) + G.nz * F.nz;
            const float perp2 = fmaxf(ng2 - d * (d / nf2), 0.0f);
            drop = ((d * q - rho * sqrtf(perp2)) - G.w) > mr * sqrtf(ng2);
          }
        }
      }
    }
    // this round's faces in list order: kept ones after the kept so far,
    // dropped ones from the end of the range backwards
    const uint32_t oct = (uint32_t)(ballot(f < a1 && !drop) >> shift) & 0xffu;
    const uint32_t octd = (uint32_t)(ballot(f < a1 && drop) >> shift) & 0xffu;
    const uint32_t below = (1u << j) - 1u;
    if (f < a1) {
      const int at = drop ? a1 - 1 - (dropped + __popc(octd & below)) : a0 + kept + __popc(oct & below);
      adj_out[at] = adjacency[f];
      if (diff) *reinterpret_cast<uint2*>(diff_out + 4 * (size_t)at) = *reinterpret_cast<const uint2*>(diff + 4 * (size_t)f);
    }
    kept += __popc(oct);
    dropped += __popc(octd);
  }
  if (j == 0) count[s] = kept;
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_prune_faces(const float* spheres, int32_t n_sites, const int32_t* adjacency, const int32_t* adjacency_offsets,
                   int64_t n_edges, const uint16_t* adjacency_diff, float margin, int32_t* adjacency_out,
                   uint16_t* diff_out, int32_t* face_count, void* stream) {
  if (n_sites < 0 || n_edges < 0 || !(margin >= 0.0f) || (n_sites > 0 && (!spheres || !adjacency_offsets || !face_count)) ||
      (n_edges > 0 && (!adjacency || !adjacency_out || (adjacency_diff && !diff_out))))
    return set_error(PF_ERR_ARG, "pf_prune_faces: bad arguments");
  if (reinterpret_cast<uintptr_t>(spheres) % 16 ||
      (reinterpret_cast<uintptr_t>(adjacency_diff) | reinterpret_cast<uintptr_t>(diff_out)) % 8)
    return set_error(PF_ERR_ARG, "pf_prune_faces: spheres must be 16-byte and the face encodings 8-byte aligned");
  if (n_sites == 0) return PF_OK;
  const int64_t threads = (int64_t)n_sites * PRUNE_LANES;
  prune_faces_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const float4*>(spheres), n_sites, adjacency, adjacency_offsets, n_edges,
      n_edges > 0 ? adjacency_diff : nullptr, margin, adjacency_out, diff_out, face_count);
  PF_HIP(hipGetLastError());
  return PF_OK;
}

int pf_power_restrict(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                      const int32_t* adjacency_offsets, int64_t n_edges, uint8_t* keep, int32_t* counts,
                      void* stream) {
  if (n_sites < 0 || n_edges < 0 || !counts || (n_sites > 0 && (!spheres || !adjacency_offsets)) ||
      (n_edges > 0 && (!adjacency || !keep)))
    return set_error(PF_ERR_ARG, "pf_power_restrict: bad arguments");
  if (reinterpret_cast<uintptr_t>(spheres) % 16)
    return set_error(PF_ERR_ARG, "pf_power_restrict: spheres must be 16-byte aligned");
  hipStream_t st = as_stream(stream);
  const int64_t span = max(n_edges, (int64_t)n_sites + 1);
  power_fill_kernel<<<(unsigned)((span + 255) / 256), 256, 0, st>>>(keep, n_edges, counts, n_sites);
  if (n_sites > 0 && n_edges > 0) {
    const float4* sp = reinterpret_cast<const float4*>(spheres);
    power_faces_kernel<<<n_sites, POWER_THREADS, 0, st>>>(sp, n_sites, adjacency, adjacency_offsets, n_edges, keep,
                                                          counts);
    power_dead_kernel<<<(n_sites + 255) / 256, 256, 0, st>>>(sp, n_sites, adjacency, adjacency_offsets, n_edges,
                                                             keep, counts);
  }
  PF_HIP(hipGetLastError());
  return PF_OK;
}

}  // extern "C"
