"""ORACLE -- test infrastructure only. ctypes wrapper of oracle/build/liborcl.so,
the CPU restatement of the reference renderer hot path (see pf_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker; the product never imports it.
"""

import ctypes
import os
import subprocess

import numpy as np

from powerfoam_amd._abi import PfCamera, PfRasterBwdArgs, PfRasterFwdArgs, PfScene, PfTileLists, tiles_hw

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "build", "liborcl.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        lib.orc_bin_count.argtypes = [vp, vp, ctypes.c_int32, vp]
        lib.orc_bin_count.restype = ctypes.c_int64
        lib.orc_bin_lists.argtypes = [vp, vp, ctypes.c_int32, vp, vp, vp]
        lib.orc_bin_lists.restype = None
        lib.orc_site_footprint_dbg.argtypes = [vp, vp, vp, vp]
        lib.orc_site_footprint_dbg.restype = ctypes.c_int
        lib.orc_tile_obb_test_dbg.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
        lib.orc_tile_obb_test_dbg.restype = ctypes.c_int
        lib.orc_cone_tree_info.argtypes = [ctypes.c_int, ctypes.c_int, vp]
        lib.orc_cones.argtypes = [vp, vp]
        lib.orc_site_cone_tiles_dbg.argtypes = [vp, vp, vp]
        lib.orc_site_cone_tiles_dbg.restype = ctypes.c_int
        lib.orc_adj_encode.argtypes = [vp, ctypes.c_int32, vp, vp, vp]
        lib.orc_adj_encode.restype = None
        lib.orc_f32_to_f16_array.argtypes = [vp, ctypes.c_int64, vp]
        lib.orc_f16_to_f32_array.argtypes = [vp, ctypes.c_int64, vp]
        for p in ("f32", "f64"):
            getattr(lib, f"orc_trace_fwd_{p}").argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_float,
                                                           ctypes.c_float, ctypes.c_int, vp]
            getattr(lib, f"orc_trace_bwd_{p}").argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_float,
                                                           ctypes.c_float, ctypes.c_int, vp, vp]
        lib.orc_set_fragile.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        lib.orc_set_fragile.restype = None
        lib.orc_power_face_encode.argtypes = [vp, ctypes.c_int32, vp, vp, vp]
        lib.orc_power_face_encode.restype = None
        lib.orc_power_restrict.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp]
        lib.orc_power_restrict.restype = None
        lib.orc_power_face.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp, ctypes.c_int64]
        lib.orc_power_face.restype = ctypes.c_int
        for p in ("f32", "f64"):
            getattr(lib, f"orc_raster_vis_{p}").argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                                            ctypes.c_float, vp, vp]
            getattr(lib, f"orc_raster_rgb_{p}").argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_float, vp]
            getattr(lib, f"orc_raster_fwd_{p}").argtypes = [vp, vp]
            getattr(lib, f"orc_raster_bwd_{p}").argtypes = [vp, vp]
        _lib = lib
    return _lib


FRAGILE_DEFAULT = (1e-6, 0.0, 1e-5)


def set_fragile(h2_rel=1e-6, h2_abs=0.0, dt_rel=1e-5):
    """Thresholds of raster_bwd's `fragile` sites: a ray with |h^2| <
    h2_rel r^2 + h2_abs r |o - c| (near tangent to the sphere; h2_abs scales
    the misplacement a float32 ray carries) or a clipped interval within
    dt_rel (1 + t_far) of zero length."""
    load().orc_set_fragile(float(h2_rel), float(h2_abs), float(dt_rel))


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _addr(a):
    return None if a is None else a.ctypes.data


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def camera(eye, right, up, width, height, ray_maps=None):
    """PfCamera from plain sequences; ray_maps (float32 [H,W,6]) is kept alive
    by the returned tuple (cam, keepalive)."""
    c = PfCamera()
    c.eye[:] = [float(v) for v in eye]
    c.right[:] = [float(v) for v in right]
    c.up[:] = [float(v) for v in up]
    c.width = int(width)
    c.height = int(height)
    rm = None
    if ray_maps is not None:
        rm = np.ascontiguousarray(ray_maps, dtype=np.float32)
        c.ray_maps = rm.ctypes.data
    c.fov_cos_cutoff = 0.0
    return c, rm


def bin_lists(cam, spheres):
    """(offsets int64 [tiles+1], sites int32 [M], keys uint32 [M])."""
    lib = load()
    sp = _f32(spheres)
    th, tw = tiles_hw(cam.height, cam.width)
    counts = np.zeros(th * tw + 1, np.int32)
    lib.orc_bin_count(ctypes.byref(cam), _p(sp), sp.shape[0], _p(counts))
    offsets = np.cumsum(counts.astype(np.int64))
    m = int(offsets[-1])
    sites = np.zeros(max(m, 1), np.int32)
    keys = np.zeros(max(m, 1), np.uint32)
    lib.orc_bin_lists(ctypes.byref(cam), _p(sp), sp.shape[0], _p(offsets), _p(sites), _p(keys))
    return offsets, sites[:m], keys[:m]


def cone_tree_info(width, height):
    """(nodes, levels, tile_level) of the cone tree of a width x height image."""
    out = np.zeros(3, np.int32)
    load().orc_cone_tree_info(int(width), int(height), _p(out))
    return int(out[0]), int(out[1]), int(out[2])


def cones(cam):
    """Cone tree [nodes, 4] of a camera with ray maps (rasterize.py:468-569)."""
    nodes, _, _ = cone_tree_info(cam.width, cam.height)
    out = np.zeros((nodes, 4), np.float32)
    load().orc_cones(ctypes.byref(cam), _p(out))
    return out


def site_cone_tiles(cam, sphere):
    """Tiles the cone-tree descent bins one site to, or None if culled."""
    th, tw = tiles_hw(cam.height, cam.width)
    out = np.zeros(th * tw, np.int32)
    n = load().orc_site_cone_tiles_dbg(ctypes.byref(cam), _p(_f32(sphere)), _p(out))
    return None if n < 0 else out[:n]


def site_footprint(cam, sphere):
    lib = load()
    sp = _f32(sphere)
    rng = np.zeros(4, np.int32)
    obb = np.zeros(6, np.float32)
    ok = lib.orc_site_footprint_dbg(ctypes.byref(cam), _p(sp), _p(rng), _p(obb))
    return (rng, obb) if ok else None


def tile_obb_test(cam, tile_i, tile_j, obb):
    lib = load()
    o = _f32(obb)
    return bool(lib.orc_tile_obb_test_dbg(int(tile_i), int(tile_j), ctypes.byref(cam), _p(o)))


def adj_encode(spheres, adjacency, offsets):
    lib = load()
    sp = _f32(spheres)
    adj = np.ascontiguousarray(adjacency, np.int32)
    off = np.ascontiguousarray(offsets, np.int32)
    out = np.zeros((max(adj.shape[0], 1), 4), np.uint16)
    lib.orc_adj_encode(_p(sp), sp.shape[0], _p(adj), _p(off), _p(out))
    return out[: adj.shape[0]]


def power_face_encode(spheres, adjacency, offsets):
    """(c_adj - c, pm_adj - pm) in binary16 (benchmark.py:297-302)."""
    lib = load()
    sp = _f32(spheres)
    adj = np.ascontiguousarray(adjacency, np.int32)
    off = np.ascontiguousarray(offsets, np.int32)
    out = np.zeros((max(adj.shape[0], 1), 4), np.uint16)
    lib.orc_power_face_encode(_p(sp), sp.shape[0], _p(adj), _p(off), _p(out))
    return out[: adj.shape[0]]


def power_restrict(spheres, adjacency, offsets):
    """Restricted power adjacency (orc_power.h): (keep uint8 [M], counts int32
    [N+1]) over the given symmetric, ascending (Cech) lists."""
    sp = _f32(spheres)
    adj = np.ascontiguousarray(adjacency, np.int32)
    off = np.ascontiguousarray(offsets, np.int32)
    n = sp.shape[0]
    keep = np.zeros(max(adj.size, 1), np.uint8)
    counts = np.zeros(n + 1, np.int32)
    load().orc_power_restrict(_p(sp), n, _p(adj), _p(off), _p(keep), _p(counts))
    return keep[:adj.size], counts


def power_face(spheres, a, b, neighbours):
    """The edge test of power_restrict for (a, b) against a's Cech list."""
    sp = _f32(spheres)
    nb = np.ascontiguousarray(neighbours, np.int32)
    return bool(load().orc_power_face(_p(sp), int(a), int(b), _p(nb), nb.size))


def restricted_adjacency(spheres, adjacency, offsets):
    """(adjacency int32, offsets int32) of the kept entries."""
    keep, counts = power_restrict(spheres, adjacency, offsets)
    return np.asarray(adjacency)[keep.astype(bool)].astype(np.int32), np.cumsum(counts).astype(np.int32)


def live_sites(eye, spheres, adjacency, offsets, diff=None, kind=0, return_margin=False):
    """Sites that a ray from `eye` can shade (uint8, pf_live_sites): 0 where the
    sphere lies beyond the plane of one of its faces by more than the
    compositing kernels' rounding margin. Not a reference function: the
    reference bins every site; this predicate only says which of them the
    compositing can skip. The faces are read as the kernels read them:
    `diff` (uint16 binary16 bits [E,4]) in the radius form (kind 0,
    adj_encode) or the power-weight form (kind 1, power_face_encode), or the
    full-precision neighbour spheres when diff is None. float32 throughout.
    With return_margin, also the per-edge decision value (beyond - margin,
    in units of |n|; > 0 culls) for borderline checks."""
    sp = np.asarray(spheres, np.float32)
    adj = np.asarray(adjacency, np.int64)
    off = np.asarray(offsets, np.int64)
    n = sp.shape[0]
    deg = np.diff(off)
    src = np.repeat(np.arange(n), deg)
    c = sp[src, :3]
    r = sp[src, 3]
    eye = np.asarray(eye, np.float32).reshape(3)
    if diff is not None:
        d = np.asarray(diff, np.uint16).view(np.float16).astype(np.float32).reshape(-1, 4)
        nv, q = d[:, :3], d[:, 3]
        if kind == 1:
            h = q - (c * nv).sum(1, dtype=np.float32)
        else:
            h = np.float32(0.5) * ((nv * nv).sum(1, dtype=np.float32) + r * r - q * q)
    else:
        a = sp[adj]
        nv = a[:, :3] - c
        h = np.float32(0.5) * ((nv * nv).sum(1, dtype=np.float32) + r * r - a[:, 3] * a[:, 3])
    nl = np.sqrt((nv * nv).sum(1, dtype=np.float32))
    dist = np.sqrt(((c - eye) ** 2).sum(1, dtype=np.float32)) + r
    m = (np.float32(8e-3) * np.sqrt(r * dist)
         + np.float32(1e-5) * (dist + np.sqrt((c * c).sum(1, dtype=np.float32)) + np.sqrt((eye * eye).sum())))
    val = (-h - r * nl) - (m * nl + np.float32(1e-5) * np.abs(h))
    dead_e = (nl > 0) & (val > 0)
    live = np.ones(n, np.uint8)
    live[src[dead_e]] = 0
    if return_margin:
        return live, val / np.maximum(nl, np.float32(1e-30)), src
    return live


def site_bounds(eye, spheres, adjacency, offsets, diff=None, kind=0, live=None, passes=3):
    """A ball (centre, radius) per site enclosing every point the compositing
    kernels can shade for the site (pf_live_sites' bounds): the sphere,
    inflated by 2e-3 sqrt(r |c - eye|) for the rounding of tangent rays, is
    replaced, for each face whose plane (moved outward by the plane-rounding
    margin) leaves the ball's centre outside, by the smallest ball enclosing
    the cap on the cell's side; `passes` sweeps over the faces in list order.
    Faces are read as live_sites reads them. Returns (bounds [N,4] float32,
    empty [N] bool: a cap vanished, so nothing of the site can shade). Not a
    reference function: the reference bins by the sphere alone."""
    sp = np.asarray(spheres, np.float32)
    adj = np.asarray(adjacency, np.int64)
    off = np.asarray(offsets, np.int64)
    n = sp.shape[0]
    eye = np.asarray(eye, np.float32).reshape(3)
    deg = np.diff(off)
    src = np.repeat(np.arange(n), deg)
    c = sp[src, :3]
    r = sp[src, 3]
    if diff is not None:
        d = np.asarray(diff, np.uint16).view(np.float16).astype(np.float32).reshape(-1, 4)
        nv, q = d[:, :3], d[:, 3]
        h = q - (c * nv).sum(1, dtype=np.float32) if kind == 1 else \
            np.float32(0.5) * ((nv * nv).sum(1, dtype=np.float32) + r * r - q * q)
    else:
        a = sp[adj]
        nv = a[:, :3] - c
        h = np.float32(0.5) * ((nv * nv).sum(1, dtype=np.float32) + r * r - a[:, 3] * a[:, 3])
    nl = np.sqrt((nv * nv).sum(1, dtype=np.float32))
    ok = nl > 0
    inv = np.where(ok, np.float32(1) / np.where(ok, nl, 1), 0).astype(np.float32)
    u = nv * inv[:, None]
    dist = np.sqrt(((c - eye) ** 2).sum(1, dtype=np.float32)) + r
    m = np.float32(1e-5) * (dist + np.sqrt((c * c).sum(1, dtype=np.float32)) + np.sqrt((eye * eye).sum())
                            + np.abs(h) * inv)
    hd = h * inv + m                      # plane offset from the site centre, moved outward
    ctr = sp[:, :3].astype(np.float32).copy()
    D = np.sqrt(((sp[:, :3] - eye) ** 2).sum(1, dtype=np.float32)) + sp[:, 3]
    rad = (sp[:, 3] + np.float32(2e-3) * np.sqrt(sp[:, 3] * D)).astype(np.float32)
    empty = np.zeros(n, bool)
    pos = np.arange(adj.size) - off[src]
    for _ in range(passes):
        for k in range(int(deg.max()) if deg.size else 0):
            e = np.nonzero((pos == k) & ok)[0]
            s_ = src[e]
            # signed distance of the ball centre inside the half-space y.u <= hd
            delta = hd[e] - ((ctr[s_] - c[e]) * u[e]).sum(1, dtype=np.float32)
            gone = delta <= -rad[s_]
            empty[s_[gone]] = True
            cut = (delta < 0) & ~gone
            sc, ec, dc = s_[cut], e[cut], delta[cut]
            ctr[sc] = ctr[sc] + u[ec] * dc[:, None]
            rad[sc] = np.sqrt(np.maximum(rad[sc] * rad[sc] - dc * dc, 0)).astype(np.float32)
    bounds = np.concatenate([ctr, rad[:, None]], 1).astype(np.float32)
    if live is not None:
        empty |= np.asarray(live) == 0
    return bounds, empty


def restrict_lists(cam, eye, right, up, offsets, sites, bounds, live):
    """Tile lists as binning with pf_live_sites' mask and balls makes them:
    the entries of live sites whose tile also meets the ball's footprint (a
    ball that fails the sphere cull -- closer than 4 radii to the eye, or
    behind it -- does not restrict its site), order kept."""
    b = np.asarray(bounds, np.float32)
    eye = np.asarray(eye, np.float32)
    fwd = np.cross(np.asarray(up, np.float32), np.asarray(right, np.float32)).astype(np.float32)
    v = b[:, :3] - eye
    free = (np.sqrt((v * v).sum(1, dtype=np.float32)) < np.float32(4) * b[:, 3]) | \
        ((v * fwd).sum(1, dtype=np.float32) < np.float32(0.1))
    ob, sb, _ = bin_lists(cam, np.where(free[:, None], np.float32(1e6), b).astype(np.float32))
    off = np.asarray(offsets, np.int64)
    tile = np.repeat(np.arange(off.size - 1), np.diff(off))
    tb = np.repeat(np.arange(ob.size - 1), np.diff(ob))
    n = b.shape[0]
    in_ball = np.isin(tile * np.int64(n) + sites, tb * np.int64(n) + sb)
    keep = (np.asarray(live)[sites] != 0) & (in_ball | free[sites])
    cnt = np.bincount(tile[keep], minlength=off.size - 1)
    out = np.zeros_like(off)
    out[1:] = np.cumsum(cnt)
    return out, np.asarray(sites)[keep]


def filter_lists(offsets, sites, live):
    """Tile lists with the sites live marks 0 left out (order kept)."""
    off = np.asarray(offsets, np.int64)
    keep = np.asarray(live)[sites] != 0
    tile = np.repeat(np.arange(off.size - 1), np.diff(off))
    cnt = np.bincount(tile[keep], minlength=off.size - 1)
    out = np.zeros_like(off)
    out[1:] = np.cumsum(cnt)
    return out, np.asarray(sites)[keep]


def raster_rgb(cam, inp, offsets, sites, threshold=1e-3, precision="f64"):
    """Benchmark forward (colour only); inp.adjacency_diff holds the power-face encoding."""
    lib = load()
    dt = np.float64 if precision == "f64" else np.float32
    offsets = np.ascontiguousarray(offsets, np.int64)
    sites = np.ascontiguousarray(sites, np.int32)
    if sites.size == 0:
        sites = np.zeros(1, np.int32)
    color = np.zeros((cam.height, cam.width, 3), dt)
    sc = inp.struct()
    getattr(lib, f"orc_raster_rgb_{precision}")(ctypes.byref(cam), ctypes.byref(sc), _p(offsets), _p(sites),
                                                offsets.shape[0] - 1, threshold, _p(color))
    return color


def f32_to_f16_bits(x):
    lib = load()
    x = _f32(x).ravel()
    y = np.zeros(x.shape[0], np.uint16)
    lib.orc_f32_to_f16_array(_p(x), x.shape[0], _p(y))
    return y


class _VisOut(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("color", "depth", "normal", "alpha", "inter")]


def raster_vis(cam, inp, offsets, sites, threshold=1e-3, max_intersections=1024, depth_quantile=0.0,
               bkgd_color=(0.0, 0.0, 0.0), precision="f64"):
    """visualization_kernel (rasterize.py:739-890): dict of color, depth,
    normal, alpha, intersections."""
    lib = load()
    dt = np.float64 if precision == "f64" else np.float32
    H, W = cam.height, cam.width
    offsets = np.ascontiguousarray(offsets, np.int64)
    sites = np.ascontiguousarray(sites, np.int32)
    if sites.size == 0:
        sites = np.zeros(1, np.int32)
    out = dict(color=np.zeros((H, W, 3), dt), depth=np.zeros((H, W), dt), normal=np.zeros((H, W, 3), dt),
               alpha=np.zeros((H, W), dt), intersections=np.zeros((H, W), np.int32))
    o = _VisOut(*[_addr(out[k]) for k in ("color", "depth", "normal", "alpha", "intersections")])
    bk = np.ascontiguousarray(bkgd_color, np.float32)
    sc = inp.struct()
    getattr(lib, f"orc_raster_vis_{precision}")(ctypes.byref(cam), ctypes.byref(sc), _p(offsets), _p(sites),
                                                offsets.shape[0] - 1, threshold, int(max_intersections),
                                                depth_quantile, _p(bk), ctypes.byref(o))
    return out


class Inputs:
    """Contiguous float32/int32 copies of a scene, as the oracle reads them."""

    def __init__(self, spheres, nsigmas, texel_sites, texel_rgbh, adjacency, adjacency_offsets,
                 adjacency_diff):
        self.spheres = _f32(spheres)
        self.nsigmas = _f32(nsigmas)
        self.texel_sites = _f32(texel_sites)
        self.texel_rgbh = _f32(texel_rgbh)
        self.adjacency = np.ascontiguousarray(adjacency, np.int32)
        self.adjacency_offsets = np.ascontiguousarray(adjacency_offsets, np.int32)
        self.adjacency_diff = None if adjacency_diff is None else np.ascontiguousarray(adjacency_diff, np.uint16)
        self.n = self.spheres.shape[0]
        self.T = self.texel_sites.shape[1]

    def struct(self):
        s = PfScene()
        s.n_sites = self.n
        s.n_texels = self.T
        s.spheres = _addr(self.spheres)
        s.nsigmas = _addr(self.nsigmas)
        s.texel_sites = _addr(self.texel_sites)
        s.texel_rgbh = _addr(self.texel_rgbh)
        s.adjacency = _addr(self.adjacency) if self.adjacency.size else None
        s.adjacency_offsets = _addr(self.adjacency_offsets)
        s.adjacency_diff = _addr(self.adjacency_diff) if self.adjacency_diff is not None and self.adjacency_diff.size else None
        return s


class _FwdOut(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in
                ("tile_early_stop", "color", "log_t", "normal_distance", "normal", "quantile_depths",
                 "err", "contrib", "point_err")]


class _BwdOut(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in
                ("grad_spheres", "grad_nsigmas", "grad_texel_sites", "grad_texel_rgb", "grad_texel_height",
                 "fragile")]


def raster_fwd(cam, inp, offsets, sites, depth_quantiles=None, ray_gt=None, threshold=1e-3,
               return_point_err=False, loss_accumulation=False, precision="f64"):
    lib = load()
    dt = np.float64 if precision == "f64" else np.float32
    H, W = cam.height, cam.width
    n_tiles = offsets.shape[0] - 1
    offsets = np.ascontiguousarray(offsets, np.int64)
    sites = np.ascontiguousarray(sites, np.int32)
    if sites.size == 0:
        sites = np.zeros(1, np.int32)
    dq = _f32(depth_quantiles)
    gt = _f32(ray_gt)
    a = PfRasterFwdArgs()
    a.cam = cam
    a.scene = inp.struct()
    a.tiles = PfTileLists(n_tiles, _addr(offsets), _addr(sites))
    a.depth_quantiles = _addr(dq)
    a.n_quantiles = 0 if dq is None else dq.shape[-1]
    a.ray_gt = _addr(gt)
    a.transmittance_threshold = threshold
    a.return_point_err = int(return_point_err)
    a.loss_accumulation = int(loss_accumulation)
    out = dict(
        tile_early_stop=np.zeros(n_tiles, np.int32),
        color=np.zeros((H, W, 3), dt), log_t=np.zeros((H, W), dt),
        normal_distance=np.zeros((H, W), dt), normal=np.zeros((H, W, 3), dt),
        quantile_depths=None if dq is None else -np.ones(dq.shape, dt),
        err=np.zeros((H, W), dt) if loss_accumulation else None,
        contrib=np.zeros(inp.n, dt),
        point_err=np.zeros(inp.n, dt) if return_point_err else None,
    )
    o = _FwdOut(*[_addr(out[k]) for k, _ in _FwdOut._fields_])
    getattr(lib, f"orc_raster_fwd_{precision}")(ctypes.byref(a), ctypes.byref(o))
    return out


def raster_bwd(cam, inp, offsets, sites, tile_early_stop, log_t, grad_color, grad_log_t,
               grad_normal_distance, grad_normal, grad_contrib, depth_quantiles=None,
               grad_quantile_depths=None, exact=False, precision="f64"):
    lib = load()
    dt = np.float64 if precision == "f64" else np.float32
    n_tiles = offsets.shape[0] - 1
    offsets = np.ascontiguousarray(offsets, np.int64)
    sites = np.ascontiguousarray(sites, np.int32)
    if sites.size == 0:
        sites = np.zeros(1, np.int32)
    es = np.ascontiguousarray(tile_early_stop, np.int32)
    keep = [_f32(x) for x in (log_t, grad_color, grad_log_t, grad_normal_distance, grad_normal, grad_contrib,
                              depth_quantiles, grad_quantile_depths)]
    lt, gc, glt, gnd, gn, gcon, dq, gq = keep
    a = PfRasterBwdArgs()
    a.cam = cam
    a.scene = inp.struct()
    a.tiles = PfTileLists(n_tiles, _addr(offsets), _addr(sites))
    a.depth_quantiles = _addr(dq)
    a.n_quantiles = 0 if dq is None else dq.shape[-1]
    a.tile_early_stop = _addr(es)
    a.log_t, a.grad_color, a.grad_log_t = _addr(lt), _addr(gc), _addr(glt)
    a.grad_normal_distance, a.grad_normal = _addr(gnd), _addr(gn)
    a.grad_quantile_depths, a.grad_contrib = _addr(gq), _addr(gcon)
    a.exact_gradients = int(exact)
    N, T = inp.n, inp.T
    out = dict(grad_spheres=np.zeros((N, 4), dt), grad_nsigmas=np.zeros((N, 4), dt),
               grad_texel_sites=np.zeros((N, T, 3), dt), grad_texel_rgb=np.zeros((N, T, 3), dt),
               grad_texel_height=np.zeros((N, T), dt), fragile=np.zeros(N, np.uint8))
    o = _BwdOut(*[_addr(out[k]) for k, _ in _BwdOut._fields_])
    getattr(lib, f"orc_raster_bwd_{precision}")(ctypes.byref(a), ctypes.byref(o))
    out["fragile"] = out["fragile"].astype(bool)
    return out


class _TraceOut(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in
                ("color", "log_t", "normal", "depth", "q_step", "q_sigma", "cells", "n_cells", "fragile")] + \
               [("K", ctypes.c_int)]


class _TraceBwdIn(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in
                ("color", "normal", "log_t", "q_step", "q_sigma", "g_color", "g_normal", "g_depth", "g_opacity")]


def trace_fwd(cam, inp, start_site, start_sites=None, threshold=1e-3, depth_q=0.5, max_steps=1 << 16,
              record_cells=0, precision="f64"):
    """Ray-tracer forward; inp.adjacency_diff is the power-face encoding or None (exact).
    `fragile` [H,W] marks rays that cross an ill-conditioned cell (orc_trace_fwd)."""
    lib = load()
    dt = np.float64 if precision == "f64" else np.float32
    H, W = cam.height, cam.width
    P = H * W
    ss = None if start_sites is None else np.ascontiguousarray(start_sites, np.int32)
    out = dict(color=np.zeros((H, W, 3), dt), log_t=np.zeros((H, W), dt), normal=np.zeros((H, W, 3), dt),
               depth=np.zeros((H, W), dt), q_step=np.zeros((H, W), np.int32), q_sigma=np.zeros((H, W), dt),
               cells=np.full((H, W, max(record_cells, 1)), -1, np.int32) if record_cells else None,
               n_cells=np.zeros((H, W), np.int32), fragile=np.zeros((H, W), np.uint8))
    o = _TraceOut(*[_addr(out[k]) for k, _ in _TraceOut._fields_[:-1]], int(record_cells))
    sc = inp.struct()
    getattr(lib, f"orc_trace_fwd_{precision}")(ctypes.byref(cam), ctypes.byref(sc), _p(ss), int(start_site),
                                               threshold, depth_q, max_steps, ctypes.byref(o))
    out["fragile"] = out["fragile"].astype(bool)
    return out


def trace_bwd(cam, inp, start_site, fwd, g_color, g_normal, g_depth, g_opacity, start_sites=None, threshold=1e-3,
              depth_q=0.5, max_steps=1 << 16, precision="f64"):
    """Exact adjoint of trace_fwd (front to back); `fwd` must come from trace_fwd
    at the same precision."""
    lib = load()
    dt = np.float64 if precision == "f64" else np.float32
    ss = None if start_sites is None else np.ascontiguousarray(start_sites, np.int32)
    keep = dict(color=np.ascontiguousarray(fwd["color"], dt),
                normal=np.ascontiguousarray(fwd["normal"], dt),
                log_t=np.ascontiguousarray(fwd["log_t"], dt),
                q_step=np.ascontiguousarray(fwd["q_step"], np.int32),
                q_sigma=np.ascontiguousarray(fwd["q_sigma"], dt),
                g_color=_f32(g_color), g_normal=_f32(g_normal), g_depth=_f32(g_depth), g_opacity=_f32(g_opacity))
    bi = _TraceBwdIn(*[_addr(keep[k]) for k, _ in _TraceBwdIn._fields_])
    N, T = inp.n, inp.T
    out = dict(grad_spheres=np.zeros((N, 4), dt), grad_nsigmas=np.zeros((N, 4), dt),
               grad_texel_sites=np.zeros((N, T, 3), dt), grad_texel_rgb=np.zeros((N, T, 3), dt),
               grad_texel_height=np.zeros((N, T), dt), fragile=None)
    o = _BwdOut(*[_addr(out[k]) for k, _ in _BwdOut._fields_])
    sc = inp.struct()
    getattr(lib, f"orc_trace_bwd_{precision}")(ctypes.byref(cam), ctypes.byref(sc), _p(ss), int(start_site),
                                               threshold, depth_q, max_steps, ctypes.byref(bi), ctypes.byref(o))
    del out["fragile"]
    return out
