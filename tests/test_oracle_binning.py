"""Pins of the oracle's tile binning (the FP32 mirror the GPU is compared with
bit-exact), by checks independent of its formulas:

- coverage: every pixel whose ray (FP64) passes through a site's ball lies in
  a tile the site is binned to (the projected-ellipse / OBB / SAT chain must
  be conservative), for every site that is not culled;
- culling: sites behind the camera or with |c - eye| < 4r are never binned;
- order: each tile's list ascends by |c - eye|^2 - r^2 (then site id), and the
  counts equal the list lengths.
A dropped term or wrong sign in the conic, the axis or the pixel mapping
breaks coverage.
"""

import numpy as np
import pytest

from oracle import orc
from powerfoam_amd import synth
from powerfoam_amd.camera import look_at


def _cam(eye, target, W, H, fov):
    t = look_at(eye, target, W, H, fov)
    c, _ = orc.camera(t.eye.numpy(), t.right.numpy(), t.up.numpy(), W, H)
    return t, c


def _pixel_rays(t):
    H, W = t.height, t.width
    right, up = t.right.numpy().astype(np.float64), t.up.numpy().astype(np.float64)
    fwd = np.cross(up, right)
    fwd /= np.linalg.norm(fwd)
    ii, jj = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    x = 2.0 * jj / (W - 1) - 1.0
    y = 1.0 - 2.0 * ii / (H - 1)
    d = x[..., None] * right + y[..., None] * up + fwd
    return d / np.linalg.norm(d, axis=-1, keepdims=True), ii, jj


@pytest.mark.parametrize("seed,eye,fov", [(0, (0.0, 0.0, 3.0), 60.0), (1, (2.5, 1.0, 1.5), 75.0),
                                          (2, (0.2, -0.3, 2.2), 50.0)])
def test_binning_covers_every_hit_pixel(seed, eye, fov):
    W, H = 90, 61
    t, cam = _cam(eye, (0.0, 0.0, 0.0), W, H, fov)
    s = synth.scene_uniform(seed, n=600, r_range=(0.01, 0.25), extent=1.5)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    offsets, sites, keys = orc.bin_lists(cam, sp)
    th = 1 + (H - 1) // 8
    tiles_of = [set() for _ in range(sp.shape[0])]
    for tile in range(offsets.shape[0] - 1):
        for site in sites[offsets[tile]:offsets[tile + 1]]:
            tiles_of[site].add(tile)
    d, ii, jj = _pixel_rays(t)
    e = t.eye.numpy().astype(np.float64)
    fwd_un = np.cross(t.up.numpy().astype(np.float64), t.right.numpy().astype(np.float64))
    checked = 0
    for k in range(sp.shape[0]):
        c = sp[k, :3].astype(np.float64)
        r = float(sp[k, 3])
        v = c - e
        culled = np.linalg.norm(v) < 4 * r or v @ fwd_un < 0.1
        if culled:
            assert not tiles_of[k], k
            continue
        # distance from the centre to each pixel ray (ray starts at the eye)
        along = d @ v
        perp2 = v @ v - along ** 2
        hit = (perp2 < (r * (1 - 1e-4)) ** 2) & (along > 0)
        for i, j in zip(ii[hit], jj[hit]):
            tile = (i // 8) + (j // 8) * th
            assert tile in tiles_of[k], (k, i, j)
            checked += 1
    assert checked > 1000


def test_lists_sorted_and_counts_consistent():
    t, cam = _cam((0.3, 0.4, 2.5), (0.0, 0.0, 0.0), 131, 77, 65.0)
    s = synth.scene_uniform(4, n=3000, r_range=(0.005, 0.2), extent=1.5)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    offsets, sites, keys = orc.bin_lists(cam, sp)
    assert offsets[0] == 0 and np.all(np.diff(offsets) >= 0) and offsets[-1] == sites.size
    # key is the float32 bit pattern of |c - eye|^2 - r^2 (positive for binned sites)
    v = sp[:, :3] - t.eye.numpy()[None]
    pw = (v.astype(np.float64) ** 2).sum(1) - sp[:, 3].astype(np.float64) ** 2
    assert np.all(pw[sites] > 0)
    np.testing.assert_allclose(keys.view(np.float32), pw[sites], rtol=1e-5)
    for tile in range(offsets.shape[0] - 1):
        a, b = offsets[tile], offsets[tile + 1]
        k, st = keys[a:b].astype(np.int64), sites[a:b].astype(np.int64)
        comp = (k << 32) | st
        assert np.all(np.diff(comp) > 0)


def test_empty_scene_and_all_culled():
    t, cam = _cam((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 64, 48, 60.0)
    sp = np.array([[0.0, 0.0, 5.0, 0.1], [0.0, 0.0, 2.9, 0.5]], np.float32)
    offsets, sites, _ = orc.bin_lists(cam, sp)
    assert offsets[-1] == 0 and sites.size == 0
