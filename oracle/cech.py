"""ORACLE -- test infrastructure only: numpy float32 restatement of the
reference's Cech adjacency (powerfoam/bvh.py), for checking the GPU builder.

  box      lo = c - r, hi = c + r (float32)              bvh.py:76-82
  query    boxes of i and j overlap, touching counts     bvh.py:20, 54 (wp.bvh_query_aabb)
  centre   c = 0.5 (lo + hi), r = 0.5 (hi.x - lo.x)      bvh.py:17-18, 29-30
  accept   j != i and |c_i - c_j| < r_i + r_j            bvh.py:24, 31-32
  lists    counts -> cumsum -> per-site lists            bvh.py:96-123

|v| is sqrt((x^2 + y^2) + z^2) in float32, each operation rounded; the
reference's compiler may contract it into FMAs, which moves only pairs
within a few ulps of touching. Neighbour ids are ascending per site (the
reference's order is its BVH's traversal order, which Warp leaves open).

Pinned by tests/test_oracle_cech.py: the kd-tree candidate search against an
all-pairs brute force, every disagreement with the exact FP64 predicate a
near-tie within the float32 rounding bound, and known answers (touching
balls, coincident centres, zero radii, self pairs, symmetry).
"""

import numpy as np

F = np.float32


def balls(points, radii):
    """(lo, hi, c, r) as the reference builds them, all float32."""
    p = np.ascontiguousarray(points, F)
    r = np.ascontiguousarray(radii, F)[:, None]
    lo = (p - r).astype(F)
    hi = (p + r).astype(F)
    c = (F(0.5) * (lo + hi)).astype(F)
    rr = (F(0.5) * (hi[:, 0] - lo[:, 0])).astype(F)
    return lo, hi, c, rr


def overlap(b, i, j):
    """The reference's predicate for index arrays i, j (float32, no FMA)."""
    lo, hi, c, r = b
    box = np.all((lo[i] <= hi[j]) & (lo[j] <= hi[i]), axis=1)
    d = (c[i] - c[j]).astype(F)
    sq = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(F) + d[:, 2] * d[:, 2]).astype(F)
    return box & (np.sqrt(sq) < (r[i] + r[j]).astype(F)) & (i != j)


def _csr(n, i, j):
    order = np.lexsort((j, i))
    i, j = i[order], j[order]
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, i + 1, 1)
    return j.astype(np.int32), np.cumsum(off).astype(np.int32)


def candidates_kdtree(points, radii):
    """Ordered pairs (i, j), i != j, whose centres lie within a bound no
    accepted pair can exceed: |c_i - c_j| < r_i + r_j <= 2 r_max, widened by
    a relative 1e-4 for float32 rounding."""
    from scipy.spatial import cKDTree
    _, _, c, r = balls(points, radii)
    if c.shape[0] < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    reach = 2.0 * float(r.max()) * (1.0 + 1e-4) + 1e-30
    pairs = cKDTree(c.astype(np.float64)).query_pairs(reach, output_type="ndarray").astype(np.int64)
    return np.concatenate([pairs[:, 0], pairs[:, 1]]), np.concatenate([pairs[:, 1], pairs[:, 0]])


def candidates_brute(n):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return i.ravel(), j.ravel()


def cech_adjacency(points, radii, brute=False):
    """(adjacency int32 [M], offsets int32 [N+1]), ids ascending per site."""
    n = np.asarray(points).shape[0]
    i, j = candidates_brute(n) if brute else candidates_kdtree(points, radii)
    b = balls(points, radii)
    keep = overlap(b, i, j) if i.size else np.zeros(0, bool)
    return _csr(n, i[keep], j[keep])


def fp64_margin(points, radii, i, j):
    """(|c_i - c_j| - (r_i + r_j), rounding bound) in FP64 on the float32
    inputs: a float32 decision can differ from the exact one only where
    |margin| <= bound."""
    p = np.asarray(points, np.float64)
    r = np.asarray(radii, np.float64)
    d = np.linalg.norm(p[i] - p[j], axis=1)
    s = r[i] + r[j]
    scale = np.abs(p[i]).max(1) + np.abs(p[j]).max(1) + s
    return d - s, 16.0 * np.finfo(np.float32).eps * scale


def neighbours(points, radii, sites):
    """Ascending neighbour ids of each site in `sites` (same predicate), from
    kd-tree ball queries: for scenes too large for the full lists."""
    from scipy.spatial import cKDTree
    b = balls(points, radii)
    c, r = b[2], b[3]
    tree = cKDTree(c.astype(np.float64))
    rmax = float(r.max()) if r.size else 0.0
    out = []
    for i in sites:
        js = np.asarray(tree.query_ball_point(c[i].astype(np.float64), (float(r[i]) + rmax) * (1 + 1e-4) + 1e-30),
                        np.int64)
        js = np.sort(js[overlap(b, np.full(js.size, i, np.int64), js)])
        out.append(js.astype(np.int32))
    return out
