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


# ---------------------------------------------------------------- ray maps
# Non-pinhole cameras bin through a cone tree over 2x2-pixel leaves
# (rasterize.py:202-291, 365-569). Pins: every node's cone contains the rays
# of all its pixels (FP64); coverage as above for an equidistant fisheye; and
# every listed tile has a pixel whose ray passes within twice the tile cone's
# half-angle of the ball (bounded over-inclusion).

def _fisheye(eye, target, size, fov=180.0):
    from powerfoam_amd.camera import fisheye_ray_maps
    rm = fisheye_ray_maps(eye, target, size, fov)
    t = look_at(eye, target, size, size, 60.0)
    c, keep = orc.camera(t.eye.numpy(), t.right.numpy(), t.up.numpy(), size, size, rm)
    return t, c, keep, rm


def _tree_nodes(W, H):
    dh, dw = (H + 1) // 2, (W + 1) // 2
    levels = []
    off = 0
    while True:
        levels.append((off, dh, dw))
        off += dh * dw
        if dh == 1 and dw == 1:
            return levels, off
        dh, dw = (dh + 1) // 2, (dw + 1) // 2


@pytest.mark.parametrize("size,fov", [(37, 180.0), (64, 120.0)])
def test_cone_tree_contains_its_pixels(size, fov):
    """Containment holds wherever the merge's cos(angle + child half-angle)
    does not pass pi; the reference's merge (rasterize.py:558-563) wraps
    there, which only the widest nodes of a >= 180 degree lens reach."""
    t, cam, keep, rm = _fisheye((0.1, 0.2, 0.4), (0.0, 0.0, -1.0), size, fov)
    cones = orc.cones(cam)
    levels, total = _tree_nodes(size, size)
    assert cones.shape[0] == total == orc.cone_tree_info(size, size)[0]
    assert orc.cone_tree_info(size, size)[2] == 2
    d = rm[..., 3:].astype(np.float64)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    wraps = np.zeros(total, bool)
    checked = 0
    for lvl, (off, dh, dw) in enumerate(levels):
        span = 2 << lvl
        for a in range(dh):
            for b in range(dw):
                cone = cones[off + a * dw + b].astype(np.float64)
                assert cone[3] <= 1.0
                if lvl > 0:
                    coff, cdh, cdw = levels[lvl - 1]
                    for ca in (2 * a, 2 * a + 1):
                        for cb in (2 * b, 2 * b + 1):
                            if ca < cdh and cb < cdw:
                                ch = cones[coff + ca * cdw + cb].astype(np.float64)
                                wraps[off + a * dw + b] |= wraps[coff + ca * cdw + cb] or \
                                    np.arccos(np.clip(ch[:3] @ cone[:3], -1, 1)) + np.arccos(ch[3]) > np.pi
                if wraps[off + a * dw + b]:
                    continue
                rays = d[a * span:(a + 1) * span, b * span:(b + 1) * span].reshape(-1, 3)
                ang = np.arccos(np.clip(rays @ cone[:3], -1, 1))
                assert ang.max() <= np.arccos(np.clip(cone[3], -1, 1)) + 1e-5, (lvl, a, b)
                checked += 1
    assert checked >= total - 3


@pytest.mark.parametrize("seed,eye", [(0, (0.1, 0.2, 0.4)), (1, (0.0, 0.0, 3.0))])
def test_cone_binning_covers_hits_and_bounds_extras(seed, eye):
    size = 48
    t, cam, keep, rm = _fisheye(eye, (0.0, 0.0, -1.0) if seed == 0 else (0.0, 0.0, 0.0), size)
    s = synth.scene_uniform(seed, n=500, r_range=(0.01, 0.2), extent=1.5)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1).astype(np.float32)
    offsets, sites, keys = orc.bin_lists(cam, sp)
    th = 1 + (size - 1) // 8
    tiles_of = [set() for _ in range(sp.shape[0])]
    for tile in range(offsets.shape[0] - 1):
        assert np.all(np.diff(keys[offsets[tile]:offsets[tile + 1]].astype(np.int64)) >= 0)
        for site in sites[offsets[tile]:offsets[tile + 1]]:
            tiles_of[site].add(tile)
    cones = orc.cones(cam)
    levels, _ = _tree_nodes(size, size)
    off2, dh2, dw2 = levels[2]
    d = rm[..., 3:].astype(np.float64)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    ii, jj = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    pix_tile = (ii // 8) + (jj // 8) * th
    e = t.eye.numpy().astype(np.float64)
    fwd_un = np.cross(t.up.numpy().astype(np.float64), t.right.numpy().astype(np.float64))
    hit_checked = extra_checked = 0
    for k in range(sp.shape[0]):
        c = sp[k, :3].astype(np.float64)
        r = float(sp[k, 3])
        v = c - e
        dist = np.linalg.norm(v)
        listed = orc.site_cone_tiles(cam, sp[k])
        if dist < 4 * r or v @ fwd_un < 0.1:
            assert listed is None and not tiles_of[k]
            continue
        assert set(listed.tolist()) == tiles_of[k]
        assert len(set(listed.tolist())) == len(listed)
        # coverage: rays that pass through the ball (clear of the rim)
        b = d @ v
        h2 = r * r - (dist * dist - b * b)
        hit = (h2 > 1e-6 * r * r) & (b > 0)
        for tile in np.unique(pix_tile[hit]):
            assert tile in tiles_of[k], (k, tile)
            hit_checked += 1
        # bounded over-inclusion
        beta = np.arcsin(min(r / dist, 1.0))
        ang = np.arccos(np.clip(d @ (v / dist), -1, 1))
        for tile in tiles_of[k]:
            ti, tj = tile % th, tile // th
            cone = cones[off2 + ti * dw2 + tj].astype(np.float64)
            half = np.arccos(np.clip(cone[3], -1, 1))
            assert ang[pix_tile == tile].min() - beta <= 2 * half + 1e-5, (k, tile)
            extra_checked += 1
    assert hit_checked > 50 and extra_checked > 100
