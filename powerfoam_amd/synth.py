"""Synthetic Power Foam scenes, cameras and adjacency for the benchmark workloads.

The reference renders trained checkpoints; here every input is generated from a
seeded numpy PCG64 stream with a fixed draw order
(positions -> radii -> density -> SH -> quaternions -> texel offsets -> heights),
and turned into the reference's rasterizer inputs the way its scene does
(powerfoam/scene.py:239-337):

  normals      = normalize(R(q) e_x)                        scene.py:242-257
  tangent frame= R(q) e_y, R(q) e_z                          scene.py:259-285
  texel_sites  = c + r (s0 t + s1 b), s ~ N(0, 0.1^2)        scene.py:135-147, 324-329
  texel_height = h r                                          scene.py:337

Adjacency is the reference's Cech complex (bvh.py:7-123: all pairs of
overlapping balls) for the small scene, and for the large scenes the power
(weighted Delaunay) adjacency from the lower hull of the lifted points with
the alpha filter |ci-cj| < ri+rj (benchmark.py:25-83). Sites whose power cell
is empty ("hidden") have no Delaunay edges; so that they render empty, as the
Cech complex would make them, each gets the face of a ball that contains it,
or, when none does, those of its 32 power-nearest sites whose radical
planes cut into it. Neighbour ids are ascending within each site.
"""

import math

import numpy as np

# sh_utils.py:4-49 real-SH constants (degree <= 3)
SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396)
SH_C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
         -0.4570457994644658, 1.445305721320277, -0.5900435899266435)


def sh_basis(d):
    """Real SH basis values Y_1..Y_15 (degree 1-3) at unit directions d [...,3]
    with the sh_utils.py constants, ordered as the coefficients sh[:, 1:16]."""
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    xx, yy, zz = x * x, y * y, z * z
    return np.stack([
        -SH_C1 * y, SH_C1 * z, -SH_C1 * x,
        SH_C2[0] * x * y, SH_C2[1] * y * z, SH_C2[2] * (2 * zz - xx - yy), SH_C2[3] * x * z, SH_C2[4] * (xx - yy),
        SH_C3[0] * y * (3 * xx - yy), SH_C3[1] * x * y * z, SH_C3[2] * y * (4 * zz - xx - yy),
        SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy), SH_C3[4] * x * (4 * zz - xx - yy),
        SH_C3[5] * z * (xx - yy), SH_C3[6] * x * (xx - 3 * yy),
    ], axis=-1)


def sh_eval(sh, d):
    """rgb = max(0, sh[0] + sum_k Y_k(d) sh[k]) -- the DC coefficient is the base
    colour. sh [...,16,3], d [...,3] unit."""
    y = sh_basis(d)
    return np.maximum(sh[..., 0, :] + np.einsum("...k,...kc->...c", y, sh[..., 1:, :]), 0.0)


def quat_frames(q):
    """(normal, tangent, bitangent) of normalised quaternions (w,x,y,z) -- scene.py:242-285."""
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    n = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1)
    t = np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1)
    b = np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)
    nz = lambda v: v / np.linalg.norm(v, axis=-1, keepdims=True)
    return nz(n), nz(t), nz(b)


def _finish_scene(rng, pts, radii, sigma, n_texels, height_std):
    n = pts.shape[0]
    sh = np.zeros((n, 16, 3), np.float32)
    sh[:, 0, :] = rng.uniform(0.0, 1.0, (n, 3))
    sh[:, 1:, :] = rng.normal(0.0, 0.05, (n, 15, 3))
    q = rng.standard_normal((n, 4))
    normals, tang, bitan = quat_frames(q)
    offs = rng.normal(0.0, 0.1, (n, n_texels, 2))
    heights = rng.normal(0.0, height_std, (n, n_texels)) if height_std > 0 else np.zeros((n, n_texels))
    off3 = (offs[..., 0:1] * radii[:, None, None]) * tang[:, None, :] + (offs[..., 1:2] * radii[:, None, None]) * bitan[:, None, :]
    texel_sites = pts[:, None, :].astype(np.float64) + off3
    return dict(
        points=pts.astype(np.float32),
        radii=radii.astype(np.float32),
        density=sigma.astype(np.float32),
        normals=normals.astype(np.float32),
        sh=sh,
        texel_sites=texel_sites.astype(np.float32),
        texel_height=(heights * radii[:, None]).astype(np.float32),
    )


def scene_uniform(seed=0, n=50_000, n_texels=8, height_std=0.0, extent=1.0,
                  r_range=(0.005, 0.05), sigma_lognormal=(0.0, 1.0)):
    """The small CPU-runnable scene: positions U[-1,1]^3, r log-uniform
    [0.005, 0.05], density lognormal(0, 1), SH-3 with DC U[0,1] and higher bands
    N(0, 0.05)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = rng.uniform(-extent, extent, (n, 3))
    radii = np.exp(rng.uniform(math.log(r_range[0]), math.log(r_range[1]), n))
    sigma = rng.lognormal(sigma_lognormal[0], sigma_lognormal[1], n)
    return _finish_scene(rng, pts, radii, sigma, n_texels, height_std)


def scene_gmm(seed=1, n=1_000_000, n_texels=8, height_std=0.0, n_comp=64, shell_frac=0.1):
    """The at-scale scene: 90% of the sites from a 64-component isotropic
    Gaussian mixture (means U[-2,2]^3, sigma U[0.05,0.4]), 10% uniformly
    directed on a shell with radius U[6,10]; r log-uniform [0.002, 0.08];
    density lognormal(0, 1.5); SH-3 as in the small scene."""
    rng = np.random.Generator(np.random.PCG64(seed))
    means = rng.uniform(-2.0, 2.0, (n_comp, 3))
    sig = rng.uniform(0.05, 0.4, n_comp)
    n_g = int(round((1.0 - shell_frac) * n))
    comp = rng.integers(0, n_comp, n_g)
    p_g = means[comp] + rng.standard_normal((n_g, 3)) * sig[comp, None]
    n_s = n - n_g
    d = rng.standard_normal((n_s, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p_s = d * rng.uniform(6.0, 10.0, n_s)[:, None]
    pts = np.concatenate([p_g, p_s], 0)
    radii = np.exp(rng.uniform(math.log(0.002), math.log(0.08), n))
    sigma = rng.lognormal(0.0, 1.5, n)
    return _finish_scene(rng, pts, radii, sigma, n_texels, height_std)


def circle_cameras(n_views=32, radius=4.0, height=1.0, width=1297, img_h=840, fov_deg=70.0, device="cpu"):
    """Pinhole cameras on a circle of the given radius and height, looking at the origin."""
    from .camera import look_at
    cams = []
    for k in range(n_views):
        th = 2.0 * math.pi * k / n_views
        eye = (radius * math.cos(th), height, radius * math.sin(th))
        cams.append(look_at(eye, (0.0, 0.0, 0.0), width, img_h, fov_deg, device=device))
    return cams


def _csr_from_pairs(n, i, j):
    """Symmetric CSR adjacency (ascending neighbour ids) from undirected pairs."""
    src = np.concatenate([i, j]).astype(np.int64)
    dst = np.concatenate([j, i]).astype(np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    keep = np.ones(src.shape[0], bool)
    keep[1:] = (src[1:] != src[:-1]) | (dst[1:] != dst[:-1])
    src, dst = src[keep], dst[keep]
    counts = np.bincount(src, minlength=n)
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    return dst.astype(np.int32), offsets.astype(np.int32)


def cech_adjacency(points, radii):
    """All pairs of overlapping balls, |ci - cj| < ri + rj (bvh.py:7-69)."""
    from scipy.spatial import cKDTree
    p = points.astype(np.float64)
    r = radii.astype(np.float64)
    tree = cKDTree(p)
    pairs = tree.query_pairs(2.0 * float(r.max()), output_type="ndarray")
    if pairs.size == 0:
        return np.zeros(0, np.int32), np.zeros(p.shape[0] + 1, np.int32)
    d = np.linalg.norm(p[pairs[:, 0]] - p[pairs[:, 1]], axis=1)
    m = d < r[pairs[:, 0]] + r[pairs[:, 1]]
    return _csr_from_pairs(p.shape[0], pairs[m, 0], pairs[m, 1])


def power_adjacency(points, radii, alpha_complex=True, hidden_cap=32, return_stats=False):
    """Power (weighted Delaunay) adjacency via the lower hull of the lifted points
    (benchmark.py:25-83), alpha-filtered, plus clipping faces for hidden sites."""
    p = points.astype(np.float64)
    r = radii.astype(np.float64)
    e, low = _power_edges(p, r)
    return _power_csr(p, r, e, low, alpha_complex, hidden_cap, return_stats)


def power_adjacency_pair(points, radii, hidden_cap=32, return_stats=False):
    """(alpha-filtered, full) power adjacency from one lower hull: the
    rasterizer's and the ray tracer's inputs for the same scene."""
    p = points.astype(np.float64)
    r = radii.astype(np.float64)
    e, low = _power_edges(p, r)
    return (_power_csr(p, r, e, low, True, hidden_cap, return_stats),
            _power_csr(p, r, e, low, False, hidden_cap, return_stats))


def _power_edges(p, r):
    """Edges of the weighted Delaunay triangulation (lower hull of the points
    lifted to |c|^2 - r^2, benchmark.py:25-50) and its tetrahedra."""
    from scipy.spatial import ConvexHull
    w = (p * p).sum(1) - r * r
    hull = ConvexHull(np.concatenate([p, w[:, None]], 1))
    low = hull.simplices[hull.equations[:, 3] < 0]
    e = np.concatenate([low[:, [0, 1]], low[:, [1, 2]], low[:, [2, 0]], low[:, [0, 3]], low[:, [1, 3]], low[:, [2, 3]]])
    return np.unique(np.sort(e, 1), axis=0), low


def _power_csr(p, r, e, low, alpha_complex, hidden_cap, return_stats):
    from scipy.spatial import cKDTree
    n = p.shape[0]
    if alpha_complex:
        d = np.linalg.norm(p[e[:, 0]] - p[e[:, 1]], axis=1)
        e = e[d < r[e[:, 0]] + r[e[:, 1]]]
    visible = np.zeros(n, bool)
    visible[low.ravel()] = True
    hidden = np.nonzero(~visible)[0]
    ei, ej = [e[:, 0]], [e[:, 1]]
    n_contained = 0
    if hidden.size:
        # power-nearest site of each hidden centre: |x-c|^2 - r^2 is the squared
        # Euclidean distance to (c, sqrt(R^2 - r^2)) in 4-D, minus R^2
        R2 = float((r * r).max())
        lift = np.concatenate([p, np.sqrt(np.maximum(R2 - r * r, 0.0))[:, None]], 1)
        tree4 = cKDTree(lift)
        q = np.concatenate([p[hidden], np.zeros((hidden.size, 1))], 1)
        _, nn = tree4.query(q, k=2, workers=-1)
        j = np.where(nn[:, 0] == hidden, nn[:, 1], nn[:, 0])
        dist = np.linalg.norm(p[j] - p[hidden], axis=1)
        contained = r[j] > dist + r[hidden]
        n_contained = int(contained.sum())
        ei.append(hidden[contained])
        ej.append(j[contained])
        rest = hidden[~contained]
        if rest.size:
            # the power-nearest sites of the centre are the ones whose radical
            # planes cut deepest; keep those that overlap the ball and cut it
            k = min(hidden_cap + 1, n)
            _, nn = tree4.query(np.concatenate([p[rest], np.zeros((rest.size, 1))], 1), k=k, workers=-1)
            hh = np.repeat(rest, k)
            nb = nn.reshape(-1).astype(np.int64)
            dd = np.linalg.norm(p[nb] - p[hh], axis=1)
            delta = (dd * dd + r[hh] ** 2 - r[nb] ** 2) / (2.0 * np.maximum(dd, 1e-30))
            keep = (nb != hh) & (dd < r[nb] + r[hh]) & (delta < r[hh])
            ei.append(hh[keep])
            ej.append(nb[keep])
    # hidden-site faces are one-directional (a hidden site does not bound anyone)
    vis_i, vis_j = ei[0], ej[0]
    src = np.concatenate([vis_i, vis_j] + ei[1:]).astype(np.int64)
    dst = np.concatenate([vis_j, vis_i] + ej[1:]).astype(np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    keep = np.ones(src.shape[0], bool)
    keep[1:] = (src[1:] != src[:-1]) | (dst[1:] != dst[:-1])
    src, dst = src[keep], dst[keep]
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=offsets[1:])
    adj, off = dst.astype(np.int32), offsets.astype(np.int32)
    if return_stats:
        return adj, off, dict(visible=int(visible.sum()), hidden=int(hidden.size), contained=n_contained,
                              edges=int(adj.size))
    return adj, off
