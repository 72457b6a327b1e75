"""Shared test harness: small scenes, a GPU runner through the drop-in
Rasterizer, and an oracle runner (FP64 unless asked otherwise) on the same
inputs. Test infrastructure only."""

import math
from dataclasses import dataclass
from types import SimpleNamespace

import numpy as np

from oracle import orc
from powerfoam_amd import synth
from powerfoam_amd.camera import TorchCamera, look_at


@dataclass
class Case:
    points: np.ndarray
    radii: np.ndarray
    density: np.ndarray
    normals: np.ndarray
    texel_sites: np.ndarray
    texel_rgb: np.ndarray
    texel_height: np.ndarray
    adjacency: np.ndarray
    adjacency_offsets: np.ndarray
    cam: TorchCamera
    gt: np.ndarray
    depth_quantiles: np.ndarray = None

    @property
    def n(self):
        return self.points.shape[0]

    @property
    def width(self):
        return self.cam.width

    @property
    def height(self):
        return self.cam.height

    def spheres(self):
        return np.concatenate([self.points, self.radii[:, None]], 1).astype(np.float32)

    def nsigmas(self):
        return np.concatenate([self.normals, self.density[:, None]], 1).astype(np.float32)

    def rgbh(self):
        return np.concatenate([self.texel_rgb, self.texel_height[..., None]], -1).astype(np.float32)

    def pf_camera(self):
        c = self.cam
        rm = None if c.ray_maps is None else c.ray_maps.cpu().numpy()
        return orc.camera(c.eye.cpu().numpy(), c.right.cpu().numpy(), c.up.cpu().numpy(), c.width, c.height, rm)


def small_case(seed=0, n=500, width=64, height=48, T=8, height_std=0.0, fov=60.0, eye=(0.0, 0.0, 3.0),
               r_range=(0.02, 0.12), quantiles=None, extent=1.0, fisheye_fov=None, density_scale=1.0):
    """Uniform scene with radii large enough that tiles hold many sites and
    faces (Cech adjacency), random per-texel colours, a random target.
    fisheye_fov: an equidistant fisheye of that field of view given as ray
    maps (width == height), binned through the cone tree."""
    s = synth.scene_uniform(seed, n=n, n_texels=T, height_std=height_std, r_range=r_range, extent=extent)
    s["density"] = (s["density"] * density_scale).astype(np.float32)
    rng = np.random.default_rng(seed + 100)
    rgb = rng.uniform(0.0, 1.0, (n, T, 3)).astype(np.float32)
    adj, off = synth.cech_adjacency(s["points"], s["radii"])
    cam = look_at(eye, (0.0, 0.0, 0.0), width, height, fov)
    if fisheye_fov is not None:
        import torch
        from powerfoam_amd.camera import fisheye_ray_maps
        assert width == height
        cam.ray_maps = torch.tensor(fisheye_ray_maps(eye, (0.0, 0.0, 0.0), width, fisheye_fov))
    gt = rng.uniform(0.0, 1.0, (height, width, 3)).astype(np.float32)
    dq = None
    if quantiles:
        dq = np.broadcast_to(np.asarray(quantiles, np.float32), (height, width, len(quantiles))).copy()
    return Case(s["points"], s["radii"], s["density"], s["normals"], s["texel_sites"], rgb, s["texel_height"],
                adj, off, cam, gt, dq)


def random_grads(case, seed=0, quantiles=False):
    rng = np.random.default_rng(seed)
    H, W = case.height, case.width
    g = dict(color=rng.normal(0, 1, (H, W, 3)), opacity=rng.normal(0, 1, (H, W)),
             normal_distance=rng.normal(0, 1, (H, W)), normal=rng.normal(0, 1, (H, W, 3)),
             contrib=rng.normal(0, 1, case.n) * H * W * 1e-2)
    if case.depth_quantiles is not None:
        g["quantile_depths"] = rng.normal(0, 1, case.depth_quantiles.shape)
    return {k: v.astype(np.float32) for k, v in g.items()}


def _args(T, is_pinhole=True):
    return SimpleNamespace(render_objective="volume", num_texel_sites=T, is_pinhole=is_pinhole,
                           disable_coop_prim_load=False, disable_coop_adj_load=False)


def run_gpu(case, dev, grads_in=None, exact_gradients=False, exact_geometry=False, return_point_err=True,
            threshold=1e-3, seg_len="default", split_backward=False, loss_outputs=None, prune_faces=None):
    """Forward (+ backward with the given output gradients) through the drop-in
    Rasterizer. Returns numpy dicts. seg_len: the Rasterizer's list-splitting
    length (None: no splitting). loss_outputs: the outputs the loss is built
    from (default all of grads_in's); the others reach the backward as None.
    prune_faces: the Rasterizer's face pruning (None: its default)."""
    import torch
    from powerfoam_amd.rasterize import Rasterizer

    T = case.texel_sites.shape[1]
    r = Rasterizer(_args(T, case.cam.ray_maps is None), dev)
    if seg_len != "default":
        r.seg_len = seg_len
        r.split_min = 0  # split every list longer than seg_len
    r.split_backward = split_backward
    if prune_faces is not None:
        r.prune_faces = prune_faces
    r.exact_gradients = exact_gradients
    r.exact_geometry = exact_geometry
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    cam = TorchCamera(case.cam.eye, case.cam.right, case.cam.up, case.width, case.height,
                      None if case.cam.ray_maps is None else case.cam.ray_maps.to(dev))
    leaves = dict(points=tt(case.points), radii=tt(case.radii), density=tt(case.density),
                  normals=tt(case.normals), texel_sites=tt(case.texel_sites), texel_rgb=tt(case.texel_rgb),
                  texel_height=tt(case.texel_height))
    for v in leaves.values():
        v.requires_grad_(grads_in is not None)
    dq = tt(case.depth_quantiles) if case.depth_quantiles is not None else None
    res = r.forward(cam, dq, leaves["points"], leaves["radii"], leaves["density"], leaves["normals"],
                    leaves["texel_sites"], leaves["texel_rgb"], leaves["texel_height"],
                    tt(case.adjacency, torch.int32), tt(case.adjacency_offsets, torch.int32), tt(case.gt),
                    return_point_err, threshold)
    names = ["color", "opacity", "normal_distance", "normal", "quantile_depths", "err", "contrib", "point_err",
             "visible"]
    out = {k: (None if v is None else v.detach().cpu().numpy()) for k, v in zip(names, res)}
    out["tile_early_stop"] = None
    grads = None
    if grads_in is not None:
        terms = []
        use = loss_outputs or ("color", "opacity", "normal_distance", "normal", "contrib", "quantile_depths")
        for k in ("color", "opacity", "normal_distance", "normal", "contrib"):
            if k in use:
                terms.append((res[names.index(k)] * tt(grads_in[k])).sum())
        if "quantile_depths" in use and "quantile_depths" in grads_in and res[4] is not None:
            terms.append((res[4] * tt(grads_in["quantile_depths"])).sum())
        torch.autograd.backward(sum(terms))
        gs = {k: v.grad.detach().cpu().numpy() for k, v in leaves.items()}
        grads = dict(grad_spheres=np.concatenate([gs["points"], gs["radii"][:, None]], 1),
                     grad_nsigmas=np.concatenate([gs["normals"], gs["density"][:, None]], 1),
                     grad_texel_sites=gs["texel_sites"], grad_texel_rgb=gs["texel_rgb"],
                     grad_texel_height=gs["texel_height"])
    torch.cuda.synchronize()
    return out, grads


def oracle_inputs(case, exact_geometry=False):
    sp = case.spheres()
    diff = None if exact_geometry else orc.adj_encode(sp, case.adjacency, case.adjacency_offsets)
    return orc.Inputs(sp, case.nsigmas(), case.texel_sites, case.rgbh(), case.adjacency, case.adjacency_offsets,
                      diff)


def run_oracle(case, grads_in=None, exact_gradients=False, exact_geometry=False, precision="f64",
               return_point_err=True, threshold=1e-3):
    cam, _keep = case.pf_camera()
    sp = case.spheres()
    offsets, sites, _ = orc.bin_lists(cam, sp)
    inp = oracle_inputs(case, exact_geometry)
    out = orc.raster_fwd(cam, inp, offsets, sites, depth_quantiles=case.depth_quantiles, ray_gt=case.gt,
                         threshold=threshold, return_point_err=return_point_err, precision=precision)
    out["opacity"] = 1.0 - np.exp(out["log_t"])
    grads = None
    if grads_in is not None:
        glt = -grads_in["opacity"].astype(np.float64) * np.exp(out["log_t"])
        grads = orc.raster_bwd(cam, inp, offsets, sites, out["tile_early_stop"], out["log_t"], grads_in["color"],
                               glt, grads_in["normal_distance"], grads_in["normal"], grads_in["contrib"],
                               depth_quantiles=case.depth_quantiles,
                               grad_quantile_depths=grads_in.get("quantile_depths"), exact=exact_gradients,
                               precision=precision)
    return out, grads


GRAD_KEYS = ("grad_spheres", "grad_nsigmas", "grad_texel_sites", "grad_texel_rgb", "grad_texel_height")


def well_conditioned(ref_grads, max_fraction=0.02):
    """Rows (sites) the oracle did not flag as ill-conditioned; asserts that
    only a small fraction is flagged."""
    frag = ref_grads["fragile"]
    assert frag.mean() <= max_fraction, f"{frag.mean():.3%} of sites flagged ill-conditioned"
    return ~frag


def rel_err(a, b):
    """||a - b|| / ||b|| over a whole gradient tensor."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / max(nb, 1e-30))


def run_gpu_benchmark(case, dev, threshold=1e-2, attr_dtype="float"):
    """Rasterizer.benchmark on the GPU; returns (colour, power-face encoding as uint16)."""
    import torch
    from powerfoam_amd.rasterize import Rasterizer, power_face_encoding

    T = case.texel_sites.shape[1]
    r = Rasterizer(_args(T, case.cam.ray_maps is None), dev, attr_dtype)
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    pts, rad = tt(case.points), tt(case.radii)
    adj, off = tt(case.adjacency, torch.int32), tt(case.adjacency_offsets, torch.int32)
    diff = power_face_encoding(pts, rad, adj, off)
    cam = TorchCamera(case.cam.eye, case.cam.right, case.cam.up, case.width, case.height,
                      None if case.cam.ray_maps is None else case.cam.ray_maps.to(dev))
    col = r.benchmark(cam, pts, rad, tt(case.density), tt(case.normals), tt(case.texel_sites),
                      tt(case.texel_rgb), tt(case.texel_height), adj, off, diff, threshold)
    torch.cuda.synchronize()
    return col.cpu().numpy(), diff.cpu().numpy().view(np.uint16)


def run_oracle_benchmark(case, diff, threshold=1e-2, attr_dtype="float"):
    cam, _keep = case.pf_camera()
    sp = case.spheres()
    q = (lambda a: a.astype(np.float16).astype(np.float32)) if attr_dtype == "half" else (lambda a: a)
    inp = orc.Inputs(sp, q(case.nsigmas()), q(case.texel_sites),
                     np.concatenate([q(case.texel_rgb), q(case.texel_height)[..., None]], -1),
                     case.adjacency, case.adjacency_offsets, diff)
    offsets, sites, _ = orc.bin_lists(cam, sp)
    return orc.raster_rgb(cam, inp, offsets, sites, threshold=threshold)
