"""ORACLE -- test infrastructure only: restatement of the reference's Morton
ordering and interpenetration regulariser (powerfoam/geometry.py).

  morton codes   float32 arithmetic of morton_code_kernel (geometry.py:18-47):
                 extent = max(upper - lower, 1e-6), n = (p - lower) / extent,
                 q = trunc(min(max(n * 2^21, 0), 2^21 - 1)), bits interleaved
                 x-major by expand_bits (geometry.py:7-15)
  morton order   stable ascending sort of the codes (radix_sort_pairs with the
                 identity as values, geometry.py:80-85)
  interpenetration  areas_i = sum_j (r_i + r_j - d)^2 over i's neighbours with
                 d = |c_i - c_j| < r_i + r_j (geometry.py:88-117), FP64, and
                 its hand-written adjoint (no direction term at d = 0, as the
                 length adjoint of the reference's autodiff)

Pinned by tests/test_oracle_geometry.py: the interleave against a bit-by-bit
loop, the codes of the box corners and of a known lattice, the order against
a per-bit lexicographic sort; the regulariser against the closed form for
two balls, torch autograd (FP64) of the same sum and central differences; a
planted sign error in the adjoint fails.
"""

import numpy as np

F = np.float32


def expand_bits(v):
    v = np.asarray(v, np.uint64) & np.uint64(0x1FFFFF)
    for shift, mask in ((32, 0x001F00000000FFFF), (16, 0x001F0000FF0000FF), (8, 0x100F00F00F00F00F),
                        (4, 0x10C30C30C30C30C3), (2, 0x1249249249249249)):
        v = (v | (v << np.uint64(shift))) & np.uint64(mask)
    return v


def morton_codes(points):
    p = np.ascontiguousarray(points, F)
    lower = p.min(0)
    upper = p.max(0)
    ext = np.maximum((upper - lower).astype(F), F(1e-6))
    res = F(2097152.0)
    norm = ((p - lower).astype(F) / ext).astype(F)
    q = np.minimum(np.maximum((norm * res).astype(F), F(0.0)), res - F(1.0)).astype(np.uint64)
    x, y, z = expand_bits(q[:, 0]), expand_bits(q[:, 1]), expand_bits(q[:, 2])
    return (x * np.uint64(4) + y * np.uint64(2) + z).astype(np.int64)


def morton_sort(points):
    return np.argsort(morton_codes(points), kind="stable").astype(np.int32)


def _edges(adjacency, offsets):
    off = np.asarray(offsets, np.int64)
    i = np.repeat(np.arange(off.size - 1), np.diff(off))
    return i, np.asarray(adjacency, np.int64)


def interpenetration(points, radii, adjacency, offsets):
    p = np.asarray(points, np.float64)
    r = np.asarray(radii, np.float64)
    i, j = _edges(adjacency, offsets)
    d = np.linalg.norm(p[i] - p[j], axis=1)
    s = r[i] + r[j]
    t = np.where(d < s, (s - d) ** 2, 0.0)
    return np.bincount(i, weights=t, minlength=p.shape[0])


def interpenetration_bwd(points, radii, adjacency, offsets, grad_areas):
    """(grad_points [N,3], grad_radii [N])."""
    p = np.asarray(points, np.float64)
    r = np.asarray(radii, np.float64)
    g = np.asarray(grad_areas, np.float64)
    n = p.shape[0]
    i, j = _edges(adjacency, offsets)
    v = p[i] - p[j]
    d = np.linalg.norm(v, axis=1)
    s = r[i] + r[j]
    c = np.where(d < s, g[i] * 2.0 * (s - d), 0.0)
    u = np.where(d[:, None] > 0, v / np.where(d > 0, d, 1.0)[:, None], 0.0)
    gp = np.zeros((n, 3))
    for k in range(3):
        gp[:, k] = np.bincount(i, weights=-c * u[:, k], minlength=n) + np.bincount(j, weights=c * u[:, k],
                                                                                  minlength=n)
    gr = np.bincount(i, weights=c, minlength=n) + np.bincount(j, weights=c, minlength=n)
    return gp, gr
