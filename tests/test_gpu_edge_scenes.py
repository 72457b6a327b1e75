"""Degenerate scenes through the drop-in rasterizer against the oracle: a
single site without faces, every site behind the camera (no list entries),
and no sites at all. Outputs within 1e-4, gradients to 1e-3 relative (and
exactly zero where the oracle's are)."""

import numpy as np
import pytest

from tests import harness

pytestmark = pytest.mark.gpu


def _subset(case, keep, shift=(0.0, 0.0, 0.0), keep_faces=False):
    keep = np.asarray(keep, np.int64)
    pts = case.points[keep] + np.asarray(shift, np.float32)
    if keep_faces:
        adj, off = case.adjacency, case.adjacency_offsets
    else:
        adj, off = np.zeros(0, np.int32), np.zeros(keep.size + 1, np.int32)
    return harness.Case(pts.astype(np.float32), case.radii[keep], case.density[keep], case.normals[keep],
                        case.texel_sites[keep] + np.asarray(shift, np.float32), case.texel_rgb[keep],
                        case.texel_height[keep], adj, off, case.cam, case.gt, case.depth_quantiles)


def _cases():
    base = harness.small_case(21, n=400, quantiles=[0.5])
    near = np.nonzero(np.linalg.norm(base.points[:, :2], axis=1) < 0.3)[0]
    big = int(near[np.argmax(base.radii[near])])
    single = _subset(base, [big])
    single.density = single.density * 20.0
    return {"single site": single,
            "all behind the camera": _subset(base, np.arange(base.n), shift=(0.0, 0.0, 8.0), keep_faces=True),
            "no sites": _subset(base, [])}


@pytest.mark.parametrize("name", ["single site", "all behind the camera", "no sites"])
def test_degenerate_scene_matches_oracle(name):
    import torch
    print(f"degenerate scene: {name}", flush=True)
    case = _cases()[name]
    g = harness.random_grads(case, seed=3)
    ref, ref_g = harness.run_oracle(case, g, precision="f64")
    out, grads = harness.run_gpu(case, torch.device("cuda"), g)
    for k in ("color", "opacity", "normal_distance", "normal", "quantile_depths", "contrib", "point_err"):
        assert out[k].shape == ref[k].shape, k
        assert np.abs(out[k] - ref[k]).max(initial=0.0) < 1e-4, k
    if name == "single site":
        assert ref["opacity"].max() > 0.1
    for k in harness.GRAD_KEYS:
        assert grads[k].shape == ref_g[k].shape, k
        if not np.any(ref_g[k]):
            assert not np.any(grads[k]), k
        else:
            assert harness.rel_err(grads[k], ref_g[k]) < 1e-3, k
