/* ORACLE (test infrastructure): restricted power adjacency, the Cech pairs
 * whose power face meets the balls, in FP64 with the operation order of the
 * GPU kernel (csrc/power.hip), so the two agree bit for bit.
 *
 * The reference's power adjacency (benchmark.py:25-83) is the weighted
 * Delaunay edge set filtered by |c_i - c_j| < r_i + r_j. A renderer only needs
 * the faces that bound a site's cell inside its own ball: on the radical plane
 * of a and b the two balls cut the same disk D (|x-c_a|^2 - r_a^2 =
 * |x-c_b|^2 - r_b^2 <= 0), and the face meets the balls when D keeps a point
 * with pow_a(x) <= pow_k(x) for every other site k. Only sites whose balls
 * overlap D can cut it, so k runs over a's Cech neighbours (bvh.py:7-69)
 * whose balls also overlap b's.
 * Every such edge is a weighted Delaunay edge that passes the overlap filter,
 * and the cell of a inside its ball is the same with and without the others.
 *
 * Per canonical edge (a < b), both directed entries get the same answer:
 *   d_b = c_b - c_a, L = |d_b|, u = d_b / L, s = (L^2 + r_a^2 - r_b^2) / 2L,
 *   rho^2 = r_a^2 - s^2 (the disk; none when <= 0 or L = 0),
 *   in-plane basis e1 = normalize(u x t) (t the axis least aligned with u),
 *   e2 = u x e1; site k gives the half-plane a1 y1 + a2 y2 <= beta with
 *   d = c_k - c_a, a1 = e1.d, a2 = e2.d, beta = (|d|^2 + r_a^2 - r_k^2)/2 - s u.d.
 * Only sites whose balls overlap b can cut the disk; the others are skipped.
 * Feasibility of disk + half-planes by Seidel's incremental LP (objective:
 * smallest y1; constraints in list order; a violated constraint moves the
 * optimum onto its line, a chord of the disk clipped by the earlier ones).
 * A site left with no face is checked at its centre: if some neighbour has
 * pow_k(c_a) < -r_a^2 its cell misses its ball, and it keeps (one way) every
 * Cech neighbour whose radical plane cuts into its ball, so that it renders
 * empty as it does with the Cech adjacency.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct {
  double u[3], e1[3], e2[3], s, rho2;
} orc_pw_edge;

static inline double orc_pw_dot(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* Plane frame of edge (a, b); 0 when the disk is empty. */
static int orc_pw_frame(const float* sa, const float* sb, orc_pw_edge* f) {
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = (double)sb[k] - (double)sa[k];
  const double L2 = orc_pw_dot(d, d);
  if (!(L2 > 0.0)) return 0;
  const double L = sqrt(L2);
  const double ra2 = (double)sa[3] * (double)sa[3], rb2 = (double)sb[3] * (double)sb[3];
  const double s = ((L2 + ra2) - rb2) / (2.0 * L);
  const double rho2 = ra2 - s * s;
  if (!(rho2 > 0.0)) return 0;
  for (int k = 0; k < 3; k++) f->u[k] = d[k] / L;
  const double ax = fabs(f->u[0]), ay = fabs(f->u[1]), az = fabs(f->u[2]);
  double t[3] = {0.0, 0.0, 0.0};
  if (ax <= ay && ax <= az) t[0] = 1.0;
  else if (ay <= az) t[1] = 1.0;
  else t[2] = 1.0;
  double e1[3] = {f->u[1] * t[2] - f->u[2] * t[1], f->u[2] * t[0] - f->u[0] * t[2], f->u[0] * t[1] - f->u[1] * t[0]};
  const double n1 = sqrt(orc_pw_dot(e1, e1));
  for (int k = 0; k < 3; k++) f->e1[k] = e1[k] / n1;
  f->e2[0] = f->u[1] * f->e1[2] - f->u[2] * f->e1[1];
  f->e2[1] = f->u[2] * f->e1[0] - f->u[0] * f->e1[2];
  f->e2[2] = f->u[0] * f->e1[1] - f->u[1] * f->e1[0];
  f->s = s;
  f->rho2 = rho2;
  return 1;
}

/* d = c_k - c_a, w = (|d|^2 + r_a^2 - r_k^2) / 2 */
static inline void orc_pw_site(const float* sa, const float* sk, double* d, double* w) {
  for (int k = 0; k < 3; k++) d[k] = (double)sk[k] - (double)sa[k];
  const double ra2 = (double)sa[3] * (double)sa[3], rk2 = (double)sk[3] * (double)sk[3];
  *w = ((orc_pw_dot(d, d) + ra2) - rk2) * 0.5;
}

static inline void orc_pw_halfplane(const orc_pw_edge* f, const double* d, double w, double* a1, double* a2,
                                    double* beta) {
  *a1 = orc_pw_dot(f->e1, d);
  *a2 = orc_pw_dot(f->e2, d);
  *beta = w - f->s * orc_pw_dot(f->u, d);
}

/* Can site k (one of a's neighbours) cut the disk of (a, b)? The disk lies in
 * ball b, so only a ball that overlaps b can; the margin keeps every such
 * ball despite rounding. d = c_k - c_a, db = c_b - c_a. */
static inline int orc_pw_reaches(const double* d, const double* db, double rk, double rb) {
  const double e[3] = {d[0] - db[0], d[1] - db[1], d[2] - db[2]};
  const double s = rk + rb;
  return orc_pw_dot(e, e) < (s * s) * (1.0 + 1e-12);
}

/* Does the face of (a, b) meet the balls? The constraints come from a's
 * neighbour list nb (m sites) in order, skipping b and the sites that cannot
 * reach the disk. */
static int orc_pw_face(const float* spheres, int32_t a, int32_t b, const int32_t* nb, int64_t m) {
  const float* sa = spheres + 4 * (int64_t)a;
  const float* sb = spheres + 4 * (int64_t)b;
  orc_pw_edge f;
  if (!orc_pw_frame(sa, sb, &f)) return 0;
  const double db[3] = {(double)sb[0] - (double)sa[0], (double)sb[1] - (double)sa[1], (double)sb[2] - (double)sa[2]};
  const double rb = (double)sb[3];
  double y1 = -sqrt(f.rho2), y2 = 0.0;
  for (int64_t q = 0; q < m; q++) {
    const int32_t k = nb[q];
    if (k == b) continue;
    double d[3], w, a1, a2, beta;
    orc_pw_site(sa, spheres + 4 * (int64_t)k, d, &w);
    if (!orc_pw_reaches(d, db, (double)spheres[4 * (int64_t)k + 3], rb)) continue;
    orc_pw_halfplane(&f, d, w, &a1, &a2, &beta);
    if (a1 * y1 + a2 * y2 <= beta) continue;
    const double nn = a1 * a1 + a2 * a2;
    if (!(nn > 0.0)) return 0;
    const double x1 = a1 * beta / nn, x2 = a2 * beta / nn;
    const double h2 = f.rho2 - (x1 * x1 + x2 * x2);
    if (h2 < 0.0) return 0;
    const double h = sqrt(h2), inv = 1.0 / sqrt(nn);
    const double t1 = -a2 * inv, t2 = a1 * inv;
    double lo = -h, hi = h;
    for (int64_t p = 0; p < q; p++) {
      const int32_t kk = nb[p];
      if (kk == b) continue;
      double dd[3], ww, b1, b2, bb;
      orc_pw_site(sa, spheres + 4 * (int64_t)kk, dd, &ww);
      if (!orc_pw_reaches(dd, db, (double)spheres[4 * (int64_t)kk + 3], rb)) continue;
      orc_pw_halfplane(&f, dd, ww, &b1, &b2, &bb);
      const double den = b1 * t1 + b2 * t2;
      const double num = bb - (b1 * x1 + b2 * x2);
      if (den > 0.0) {
        const double v = num / den;
        if (v < hi) hi = v;
      } else if (den < 0.0) {
        const double v = num / den;
        if (v > lo) lo = v;
      } else if (num < 0.0) {
        return 0;
      }
      if (lo > hi) return 0;
    }
    const double t = t1 > 0.0 ? lo : hi;
    y1 = x1 + t * t1;
    y2 = x2 + t * t2;
  }
  return 1;
}

/* The edge test alone, for tests. */
int orc_power_face(const float* spheres, int32_t a, int32_t b, const int32_t* nb, int64_t m) {
  return orc_pw_face(spheres, a, b, nb, m);
}

/* keep [M] (0/1 per directed entry), counts [N+1] (counts[0] = 0, kept per
 * site). Lists must be symmetric and ascending per site. */
void orc_power_restrict(const float* spheres, int32_t n, const int32_t* adj, const int32_t* off, uint8_t* keep,
                        int32_t* counts) {
  const int64_t M = off[n];
  memset(keep, 0, (size_t)M);
  memset(counts, 0, sizeof(int32_t) * ((size_t)n + 1));
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t a = 0; a < n; a++) {
    const int64_t o0 = off[a], o1 = off[a + 1];
    for (int64_t p = o0; p < o1; p++) {
      const int32_t b = adj[p];
      if (b <= a) continue;
      if (!orc_pw_face(spheres, a, b, adj + o0, o1 - o0)) continue;
      keep[p] = 1;
      /* the reverse entry: a in b's ascending list */
      int64_t lo = off[b], hi = off[b + 1];
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (adj[mid] < a) lo = mid + 1;
        else hi = mid;
      }
      if (lo < off[b + 1] && adj[lo] == a) keep[lo] = 1;
    }
  }
  for (int32_t a = 0; a < n; a++) {
    int32_t c = 0;
    for (int64_t p = off[a]; p < off[a + 1]; p++) c += keep[p];
    counts[a + 1] = c;
  }
  /* sites left without a face */
  for (int32_t a = 0; a < n; a++) {
    if (counts[a + 1]) continue;
    const float* sa = spheres + 4 * (int64_t)a;
    const double ra = (double)sa[3];
    int dead = 0;
    for (int64_t p = off[a]; p < off[a + 1] && !dead; p++) {
      double d[3], w;
      orc_pw_site(sa, spheres + 4 * (int64_t)adj[p], d, &w);
      dead = w < 0.0;
    }
    if (!dead) continue;
    int32_t c = 0;
    for (int64_t p = off[a]; p < off[a + 1]; p++) {
      double d[3], w;
      orc_pw_site(sa, spheres + 4 * (int64_t)adj[p], d, &w);
      /* the radical plane lies at w / |d| from c_a along d: it cuts into the
       * ball when that is below r_a */
      const double L = sqrt(orc_pw_dot(d, d));
      if (L > 0.0 && w < ra * L) {
        keep[p] = 1;
        c++;
      }
    }
    counts[a + 1] = c;
  }
}
