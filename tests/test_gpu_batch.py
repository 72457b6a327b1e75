"""The training-step workload: a batch of 8 pinhole views of mixed resolution
(4x 1297x840, 2x 979x546, 2x 1920x1080) of the 3M-site scene, rasterized
forward and backward with the learned MLP colour head and the texel features
stored in BF16 (geometry and transmittance FP32), one summed L2 loss.

- BF16 storage is exact bookkeeping: the rasterizer with BF16 features on
  inputs X gives the same images, bit for bit, as the FP32 path on the
  features of X rounded to BF16, and the same gradients up to float
  summation order;
- the batch: forward of a 979x546 view against the FP64 oracle on the
  BF16-rounded features (bars of the config-scale tests), every gradient
  finite, and the MLP and geometry gradients close to those of the same
  batch with FP32 features (the rounding moves colours by at most 2^-9
  relative);
- FP64 central differences (eps 1e-4, confirmed at 1e-5 to 2e-4) on 256 sampled
  parameters of the shown sites of a sub-scene (40 sites of one radius band
  of the 3M-site scene, nearest that band's densest point, scaled about it) seen by
  the same 8-view mixed batch at 1/8 of each size,
  against the GPU's exact-mode gradients with BF16 features; the
  references' features are the BF16-rounded values.
"""

import time

import numpy as np
import pytest

from oracle import orc
from tests import configs, harness

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def test_bf16_features_equal_rounded_fp32_features():
    import torch
    case = harness.small_case(seed=6, n=800, height_std=0.05)
    gin = harness.random_grads(case, seed=2)
    out_b, g_b = _run(case, gin, "bf16")
    rounded = harness.Case(**{f: getattr(case, f) for f in harness.Case.__dataclass_fields__})
    rounded.texel_rgb = _bf16_round(case.texel_rgb)
    rounded.texel_height = _bf16_round(case.texel_height)
    assert not np.array_equal(rounded.texel_rgb, case.texel_rgb)
    out_f, g_f = _run(rounded, gin, "float")
    for k in ("color", "opacity", "normal_distance", "normal"):
        assert np.array_equal(out_b[k], out_f[k]), k
    for k in harness.GRAD_KEYS:
        assert harness.rel_err(g_b[k], g_f[k]) < 1e-6, k
    torch.cuda.synchronize()


def _run(case, gin, feature_dtype, exact=False):
    """harness.run_gpu with the rasterizer's feature storage chosen."""
    from powerfoam_amd import rasterize
    real = rasterize.Rasterizer.__init__

    def init(self, *a, **k):
        real(self, *a, **k)
        self.feature_dtype = feature_dtype

    rasterize.Rasterizer.__init__ = init
    try:
        return harness.run_gpu(case, "cuda", gin, exact_gradients=exact, exact_geometry=exact,
                               return_point_err=not exact, threshold=0.0 if exact else 1e-3)
    finally:
        rasterize.Rasterizer.__init__ = real


def _batch(rast, head, x, ts, adj, off, cams, targets):
    """Forward of every view and the summed L2 loss (as one training step
    over the batch would build it)."""
    import torch
    loss = 0.0
    outs = []
    for cam, tgt in zip(cams, targets):
        rgb = head.forward(ts, cam)
        out = rast.forward(cam, None, x["points"], x["radii"], x["density"], x["normals"], ts, rgb,
                           x["texel_height"], adj, off, tgt, False)
        loss = loss + torch.mean((out[0] - tgt) ** 2)
        outs.append(out)
    return loss, outs


def test_c4_mixed_batch_bf16_features():
    import torch
    from powerfoam_amd.rasterize import Rasterizer
    s = configs.c3_scene()
    head, _ = configs.c3_head()
    dev = "cuda"
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    x = {k: tt(s[k]).requires_grad_(True) for k in ("points", "radii", "density", "normals", "texel_height")}
    ts = tt(s["texel_sites"])
    adj, off = tt(s["adj"], torch.int32), tt(s["off"], torch.int32)
    cams = configs.c4_cameras()
    rng = np.random.default_rng(4)
    targets = [tt(rng.uniform(0, 1, (c.height, c.width, 3)).astype(np.float32)) for c in cams]
    rays = sum(c.width * c.height for c in cams)
    grads = {}
    for fd in ("float", "bf16"):
        rast = Rasterizer(harness._args(8), dev)
        rast.feature_dtype = fd
        for v in list(x.values()) + head.parameters():
            v.grad = None
        loss, _ = _batch(rast, head, x, ts, adj, off, cams, targets)  # warm-up
        loss.backward()
        for v in list(x.values()) + head.parameters():
            v.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss, outs = _batch(rast, head, x, ts, adj, off, cams, targets)
        loss.backward()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        configs.log(f"c4 batch, {fd} features: {ms:.1f} ms for 8 views ({rays / ms / 1e3:.1f} Mrays/s), "
                    f"loss {float(loss):.5f}")
        grads[fd] = {k: v.grad.detach().cpu().numpy() for k, v in x.items()}
        grads[fd].update({f"mlp_{k}": v.grad.detach().cpu().numpy() for k, v in head.params.items()})
        for k, v in grads[fd].items():
            assert np.all(np.isfinite(v)), (fd, k)
    for k in grads["float"]:
        e = harness.rel_err(grads["bf16"][k], grads["float"][k])
        configs.log(f"c4: {k} gradient, bf16 vs fp32 features: rel {e:.2e}")
        assert e < 2e-2, (k, e)
    # one 979x546 view against the FP64 oracle on the BF16-rounded features
    cam = cams[4]
    with torch.no_grad():
        rgb = head.forward(ts, cam).cpu().numpy()
    case = harness.Case(s["points"], s["radii"], s["density"], s["normals"], s["texel_sites"], _bf16_round(rgb),
                        s["texel_height"], s["adj"], s["off"], cam, np.zeros((cam.height, cam.width, 3), np.float32))
    rast = Rasterizer(harness._args(8), dev)
    rast.feature_dtype = "bf16"
    with torch.no_grad():
        out = rast.forward(cam, None, x["points"], x["radii"], x["density"], x["normals"], ts,
                           torch.tensor(rgb, device=dev), x["texel_height"], adj, off, None, False)
    ocam, _ = case.pf_camera()
    offs, sites, _ = orc.bin_lists(ocam, case.spheres())
    inp = harness.oracle_inputs(case)
    refs = {p: orc.raster_fwd(ocam, inp, offs, sites, precision=p) for p in ("f64", "f32")}
    from tests.test_gpu_configs import _out_bar
    _out_bar("c4 979x546 view, bf16 features", "color", out[0].cpu().numpy(), refs["f64"]["color"],
             refs["f32"]["color"])
    _out_bar("c4 979x546 view, bf16 features", "opacity", out[1].cpu().numpy(), 1 - np.exp(refs["f64"]["log_t"]),
             1 - np.exp(refs["f32"]["log_t"]))


def _sub_scene(n=40):
    """The n sites of radius 0.02-0.03 of the 3M-site scene nearest the densest
    point of that band (most band neighbours among 2,000 samples), scaled about that point to an
    extent of 0.6, with their Cech adjacency, random texel colours, small
    heights and densities of 0.5-5."""
    from scipy.spatial import cKDTree
    from powerfoam_amd import synth
    s = configs.c3_scene()
    rng = np.random.default_rng(12)
    # sites of one radius band, so that no sphere swallows the others' cells
    band = np.nonzero((s["radii"] > 0.02) & (s["radii"] < 0.03))[0]
    tree = cKDTree(s["points"][band])
    probe = rng.choice(band.size, 2000, replace=False)
    cnt = [len(tree.query_ball_point(s["points"][band[i]], 0.05)) for i in probe]
    centre = s["points"][band[probe[int(np.argmax(cnt))]]]
    _, j = tree.query(centre, k=n)
    idx = np.sort(band[j])
    sub = {k: np.ascontiguousarray(s[k][idx]) for k in ("points", "radii", "density", "normals", "texel_sites",
                                                        "texel_height")}
    # scaled about its centre to an extent of ~0.6 (the power diagram scales
    # with it), so that the cameras sit where the reference's cull
    # (v . (up x right) >= 0.1, un-normalised) keeps the sites; densities
    # chosen so that the cells are neither transparent nor opaque
    k = 0.6 / float(np.linalg.norm(sub["points"] - centre, axis=1).max() + sub["radii"].max())
    sub["points"] = (centre + (sub["points"] - centre) * k).astype(np.float32)
    sub["texel_sites"] = (centre + (sub["texel_sites"] - centre) * k).astype(np.float32)
    sub["radii"] = (sub["radii"] * k).astype(np.float32)
    sub["density"] = rng.uniform(0.5, 5.0, n).astype(np.float32)
    sub["texel_rgb"] = _bf16_round(rng.uniform(0, 1, (n, 8, 3)))
    sub["texel_height"] = _bf16_round(rng.normal(0, 0.05, (n, 8)) * sub["radii"][:, None])
    sub["adj"], sub["off"] = synth.cech_adjacency(sub["points"], sub["radii"])
    return sub, centre


def _sub_cameras(centre, radius, fov):
    from powerfoam_amd.camera import look_at
    sizes = [(1297, 840)] * 4 + [(979, 546)] * 2 + [(1920, 1080)] * 2
    cams = []
    for k, (w, h) in enumerate(sizes):
        th = 2.0 * np.pi * k / 8
        eye = centre + radius * np.array([np.cos(th), 0.3, np.sin(th)]) / np.sqrt(1.09)
        cams.append(look_at(eye, centre, w // 8, h // 8, fov))
    return cams


def test_c4_sub_scene_gradients_match_fp64_central_differences():
    import torch
    sub, centre = _sub_scene()
    extent = float(np.linalg.norm(sub["points"] - centre, axis=1).max() + sub["radii"].max())
    # far enough that no site is culled (|c - eye| >= 4 r), the sub-scene
    # filling about 80 % of the narrower image side
    dist = max(3.0 * extent, 5.0 * float(sub["radii"].max()) + extent)
    cams = _sub_cameras(centre.astype(np.float64), dist, float(np.degrees(2.0 * np.arctan(1.6 * extent / dist))))
    rng = np.random.default_rng(13)
    n = sub["points"].shape[0]
    wts = [dict(color=rng.normal(size=(c.height, c.width, 3)), opacity=rng.normal(size=(c.height, c.width)),
                normal=rng.normal(size=(c.height, c.width, 3)), contrib=rng.normal(size=n) * c.width * c.height * 0.05)
           for c in cams]
    print(f"c4 sub-scene: {n} sites, {sub['adj'].size} faces, extent {extent:.3f}, 8 views", flush=True)

    def case_of(cam):
        return harness.Case(sub["points"], sub["radii"], sub["density"], sub["normals"], sub["texel_sites"],
                            sub["texel_rgb"], sub["texel_height"], sub["adj"], sub["off"], cam,
                            np.zeros((cam.height, cam.width, 3), np.float32))

    def objective(contrib=None):
        L = 0.0
        for cam, w in zip(cams, wts):
            c = case_of(cam)
            ocam, _ = c.pf_camera()
            offs, sites, _ = orc.bin_lists(ocam, c.spheres())
            out = orc.raster_fwd(ocam, harness.oracle_inputs(c, exact_geometry=True), offs, sites, threshold=0.0)
            L += (out["color"] * w["color"]).sum() + ((1 - np.exp(out["log_t"])) * w["opacity"]).sum()
            L += (out["normal"] * w["normal"]).sum() + (out["contrib"] * w["contrib"]).sum()
            if contrib is not None:
                contrib += out["contrib"]
        return L

    g = None
    for cam, w in zip(cams, wts):
        gin = dict(color=w["color"], opacity=w["opacity"], normal=w["normal"], contrib=w["contrib"],
                   normal_distance=np.zeros((cam.height, cam.width)))
        gin = {k: v.astype(np.float32) for k, v in gin.items()}
        _, gv = _run(case_of(cam), gin, "bf16", exact=True)
        g = gv if g is None else {k: g[k] + gv[k] for k in g}
    torch.cuda.synchronize()
    from tests.test_oracle_raster import PARAMS
    keys = ["points", "radii", "density", "normals", "texel_rgb"]
    # parameters of the sites some view shows (the others have no gradient)
    contrib = np.zeros(n)
    objective(contrib)
    shown = np.nonzero(contrib > 1e-4 * contrib.max())[0]
    print(f"c4 sub-scene: {shown.size} of {n} sites shown", flush=True)
    assert shown.size >= 10
    tested, worst, nonzero, skipped = 0, 0.0, 0, 0
    for trial in range(640):
        if tested == 256:
            break
        k = keys[trial % len(keys)]
        arr = sub[k]
        idx = (int(rng.choice(shown)),) + tuple(int(rng.integers(0, d)) for d in arr.shape[1:])
        gk, col = PARAMS[k]
        site = idx[0]
        if k in ("points", "normals"):
            ana = g[gk][site, idx[1]]
        elif k in ("radii", "density"):
            ana = g[gk][site, col]
        else:
            ana = g[gk][idx]
        base = float(arr[idx])
        fds = []
        for eps in (1e-4, 1e-5):
            arr[idx] = base + eps
            xp = float(arr[idx])
            Lp = objective()
            arr[idx] = base - eps
            xm = float(arr[idx])
            Lm = objective()
            arr[idx] = base
            fds.append((Lp - Lm) / (xp - xm))
        # the two step sizes disagree: a constraint switches or the loss bends
        # within +-eps of this point, and neither difference is the derivative
        if abs(fds[0] - fds[1]) > 2e-4 * (abs(fds[0]) + 1.0):
            skipped += 1
            continue
        tested += 1
        nonzero += abs(fds[0]) > 1e-3
        worst = max(worst, abs(ana - fds[0]) / (abs(fds[0]) + 1.0))
        assert abs(ana - fds[0]) <= 3e-3 * abs(fds[0]) + 3e-3, (k, idx, ana, fds)
    print(f"c4 sub-scene gradcheck: {tested} parameters tested ({nonzero} with |dL/dx| > 1e-3, {skipped} skipped), "
          f"worst scaled error {worst:.2e}", flush=True)
    assert tested >= 200 and nonzero >= 100
