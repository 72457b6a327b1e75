"""The benchmark workloads at their own sizes, built once per test session.

  c0: seed 0, 50k sites uniform in [-1,1]^3, r log-uniform [0.005, 0.05],
      density lognormal(0, 1), degree-3 SH; one pinhole camera 640x480, fov
      60 deg, at (0,0,3) looking at the origin; Cech adjacency; L2 loss
      against a uniform-random target image.
  c1: seed 1, 1M sites (90% from a 64-component Gaussian mixture, 10% on a
      radius 6-10 shell), r log-uniform [0.002, 0.08], density
      lognormal(0, 1.5), degree-3 SH; 32 pinhole cameras on a radius-4
      circle at height 1, 1297x840, fov 70 deg; the power/alpha adjacency for
      the rasterizer and the full power adjacency for the ray tracer, both
      from one lower hull.
  c3: seed 2, 3M sites with the c1 distributions (the learned-head scene).

texel_rgb is the SH colour at each texel's direction from the eye
(normalize(texel_site - eye), as color_fn.py:72 takes it), evaluated in
FP64. Test infrastructure only.
"""

import functools
import time

import numpy as np

from powerfoam_amd import synth
from powerfoam_amd.camera import look_at
from tests import harness


def log(msg):
    print(f"[configs] {msg}", flush=True)


def texel_rgb(scene, eye):
    d = scene["texel_sites"].astype(np.float64) - np.asarray(eye, np.float64)[None, None]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return synth.sh_eval(scene["sh"][:, None].astype(np.float64), d).astype(np.float32)


@functools.lru_cache(maxsize=None)
def c0():
    t0 = time.time()
    s = synth.scene_uniform(0, n=50_000)
    adj, off = synth.cech_adjacency(s["points"], s["radii"])
    cam = look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 640, 480, 60.0)
    gt = np.random.default_rng(100).uniform(0.0, 1.0, (480, 640, 3)).astype(np.float32)
    case = harness.Case(s["points"], s["radii"], s["density"], s["normals"], s["texel_sites"],
                        texel_rgb(s, cam.eye.numpy()), s["texel_height"], adj, off, cam, gt,
                        np.full((480, 640, 1), 0.5, np.float32))
    log(f"c0: {case.n} sites, {adj.size} Cech edges, {time.time() - t0:.1f}s")
    return case


@functools.lru_cache(maxsize=None)
def c1_scene():
    t0 = time.time()
    s = synth.scene_gmm(1, n=1_000_000)
    (adj, off, st), (adj_f, off_f, st_f) = synth.power_adjacency_pair(s["points"], s["radii"], return_stats=True)
    s.update(adj=adj, off=off, adj_full=adj_f, off_full=off_f)
    log(f"c1 scene: {st}, full {st_f['edges']} edges, {time.time() - t0:.1f}s")
    return s


def c1_cameras():
    return synth.circle_cameras(n_views=32, width=1297, img_h=840)


@functools.lru_cache(maxsize=None)
def c1(view=0):
    s = c1_scene()
    cam = c1_cameras()[view]
    gt = np.random.default_rng(1234).uniform(0.0, 1.0, (840, 1297, 3)).astype(np.float32)
    return harness.Case(s["points"], s["radii"], s["density"], s["normals"], s["texel_sites"],
                        texel_rgb(s, cam.eye.numpy()), s["texel_height"], s["adj"], s["off"], cam, gt,
                        np.full((840, 1297, 1), 0.5, np.float32))


@functools.lru_cache(maxsize=None)
def c3_scene():
    t0 = time.time()
    s = synth.scene_gmm(2, n=3_000_000)
    adj, off, st = synth.power_adjacency(s["points"], s["radii"], return_stats=True)
    s.update(adj=adj, off=off)
    log(f"c3 scene: {st}, {time.time() - t0:.1f}s")
    return s


@functools.lru_cache(maxsize=None)
def c3_head(encoding="sh"):
    """The learned colour head on the 3M-site scene, trained on the GPU with its
    own kernels: 2000 Adam steps of 65,536 SH-labelled (site, direction)
    pairs, with the given direction encoding. Returns (head, training loss
    history)."""
    import torch
    from powerfoam_amd.mlp_head import MLPColor, fit, latents_from_sh
    s = c3_scene()
    sh = torch.tensor(s["sh"], device="cuda")
    head = MLPColor(sh.shape[0], "cuda", seed=0, latents=latents_from_sh(sh), encoding=encoding)
    t0 = time.time()
    hist = fit(head, sh, steps=2000, batch=65536, log_every=500)
    torch.cuda.synchronize()
    log(f"c3 head ({encoding}): trained in {time.time() - t0:.1f}s, loss {hist[0][1]:.4f} -> {hist[-1][1]:.4f}")
    return head, hist


def c4_cameras(scale=1):
    """The mixed-resolution batch: 4 views of 1297x840, 2 of 979x546, 2 of
    1920x1080 (each size divided by `scale`), the first 8 cameras of the
    radius-4 circle, fov 70 deg."""
    from powerfoam_amd.camera import look_at
    sizes = [(1297, 840)] * 4 + [(979, 546)] * 2 + [(1920, 1080)] * 2
    cams = []
    for k, (w, h) in enumerate(sizes):
        th = 2.0 * np.pi * k / 32
        cams.append(look_at((4.0 * np.cos(th), 1.0, 4.0 * np.sin(th)), (0.0, 0.0, 0.0), max(w // scale, 8),
                            max(h // scale, 8), 70.0))
    return cams
