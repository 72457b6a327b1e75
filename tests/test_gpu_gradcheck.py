"""GPU adjoint kernels against FP64 central finite differences.

The GPU backward runs in exact-gradient mode with full-precision neighbour
geometry (Rasterizer.exact_gradients / exact_geometry); the finite differences
(eps = 1e-4, confirmed at 1e-5) are of the FP64 oracle forward of the same
random linear functional of all outputs, on 256 sampled parameters per scene.
Samples whose two step sizes disagree (an active constraint switches within
+-eps) are skipped; at least 200 must be tested."""

import numpy as np
import pytest

from tests import harness
from tests.test_oracle_raster import PARAMS, _objective, _objective_weights, camera, make_scene

pytestmark = pytest.mark.gpu


def _gpu_grads(s, t, quantiles, wts, out_ref, dev):
    case = harness.Case(s["points"], s["radii"], s["density"], s["normals"], s["texel_sites"], s["texel_rgb"],
                        s["texel_height"], s["adj"], s["off"], t, np.zeros((t.height, t.width, 3), np.float32),
                        quantiles)
    wq = np.where(out_ref["quantile_depths"] != -1, wts["quantile_depths"], 0.0)
    gin = dict(color=wts["color"], opacity=wts["opacity"], normal_distance=wts["normal_distance"],
               normal=wts["normal"], contrib=wts["contrib"], quantile_depths=wq)
    gin = {k: v.astype(np.float32) for k, v in gin.items()}
    _, g = harness.run_gpu(case, dev, gin, exact_gradients=True, exact_geometry=True, return_point_err=False,
                           threshold=0.0)
    return g


@pytest.mark.parametrize("mode", ["uniform", "varied"])
def test_gpu_gradients_match_fp64_central_differences(mode):
    import torch
    dev = torch.device("cuda:0")
    s = make_scene(11 if mode == "uniform" else 12, texel_mode=mode, height_std=0.0 if mode == "uniform" else 0.1)
    t, cam = camera()
    H, W = t.height, t.width
    quantiles = np.broadcast_to(np.array([0.8, 0.4], np.float32), (H, W, 2)).copy()
    rng = np.random.default_rng(1)
    wts = _objective_weights(H, W, s["points"].shape[0], 2, rng)
    _, out = _objective(s, cam, wts, quantiles)
    print(f"gradcheck {mode}: {s['points'].shape[0]} sites {W}x{H}", flush=True)
    g = _gpu_grads(s, t, quantiles, wts, out, dev)
    keys = [k for k in PARAMS if not (mode == "varied" and k == "radii")]
    if mode == "uniform":
        keys = [k for k in keys if k not in ("texel_rgb", "texel_height", "texel_sites")]
    tested, worst = 0, 0.0
    for trial in range(256):
        k = keys[trial % len(keys)]
        arr = s[k]
        idx = tuple(rng.integers(0, d) for d in arr.shape)
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
            Lp, _ = _objective(s, cam, wts, quantiles)
            arr[idx] = base - eps
            xm = float(arr[idx])
            Lm, _ = _objective(s, cam, wts, quantiles)
            arr[idx] = base
            fds.append((Lp - Lm) / (xp - xm))
        if abs(fds[0] - fds[1]) > 1e-3 * (abs(fds[0]) + 1.0):
            continue
        tested += 1
        err = abs(ana - fds[0]) / (abs(fds[0]) + 1.0)
        worst = max(worst, err)
        assert abs(ana - fds[0]) <= 3e-3 * abs(fds[0]) + 3e-3, (k, idx, ana, fds)
    print(f"gradcheck {mode}: {tested} parameters tested, worst scaled error {worst:.2e}", flush=True)
    assert tested >= 200
