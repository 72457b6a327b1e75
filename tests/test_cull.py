"""Sites the tile lists may leave out (pf_live_sites, oracle.orc.live_sites).

A site whose sphere lies beyond one of its own power faces has an empty
clipped interval on every ray. The oracle's own compositing (the reference's
algorithm, which bins every site) must then give the same outputs and
gradients with and without those sites, bit for bit; and in float64 every
point of such a sphere must lie outside the site's cell. A wrong sign or a
dropped term in the predicate culls a site that shades, which the first test
sees as a changed pixel or gradient."""

import numpy as np
import pytest

from oracle import orc
from tests.harness import oracle_inputs, random_grads, small_case

GRAD_KEYS = ("grad_spheres", "grad_nsigmas", "grad_texel_sites", "grad_texel_rgb", "grad_texel_height")


def _heavy(seed=0, **kw):
    # radii up to 0.3 in the unit cube: most small spheres sit inside a
    # larger neighbour, beyond the radical plane between the two
    return small_case(seed, n=3000, r_range=(0.01, 0.3), quantiles=[0.5, 0.1], **kw)


def _lists(case, live):
    cam, _ = case.pf_camera()
    off, sites, _ = orc.bin_lists(cam, case.spheres())
    off2, sites2 = orc.filter_lists(off, sites, live)
    return cam, (off, sites), (off2, sites2)


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("exact_geometry", [False, True])
def test_oracle_forward_and_backward_unchanged_without_culled_sites(precision, exact_geometry):
    case = _heavy()
    inp = oracle_inputs(case, exact_geometry)
    live = orc.live_sites(case.cam.eye.numpy(), case.spheres(), case.adjacency, case.adjacency_offsets,
                          inp.adjacency_diff)
    assert (live == 0).mean() > 0.3
    cam, full, culled = _lists(case, live)
    assert culled[1].size < 0.8 * full[1].size
    g = random_grads(case, seed=1)
    res = []
    for off, sites in (full, culled):
        out = orc.raster_fwd(cam, inp, off, sites, depth_quantiles=case.depth_quantiles, ray_gt=case.gt,
                             return_point_err=True, precision=precision)
        glt = -g["opacity"].astype(np.float64) * np.exp(out["log_t"])
        gr = orc.raster_bwd(cam, inp, off, sites, out["tile_early_stop"], out["log_t"], g["color"], glt,
                            g["normal_distance"], g["normal"], g["contrib"], depth_quantiles=case.depth_quantiles,
                            grad_quantile_depths=g["quantile_depths"], precision=precision)
        res.append((out, gr))
    (a, ga), (b, gb) = res
    # per-pixel outputs bit for bit
    for k in ("color", "log_t", "normal_distance", "normal", "quantile_depths"):
        assert np.array_equal(a[k], b[k]), k
    # per-site sums run over tiles in an order that follows the list lengths:
    # they agree up to that summation order (a few ulp)
    rtol = 4e-6 if precision == "f32" else 1e-13
    for x, y, k in [(a[k], b[k], k) for k in ("contrib", "point_err")] + [(ga[k], gb[k], k) for k in GRAD_KEYS]:
        np.testing.assert_allclose(y, x, rtol=rtol, atol=rtol * float(np.abs(x).max()), err_msg=k)
    # the culled sites' own gradients and contributions are zero
    dead = live == 0
    assert not np.any(ga["grad_spheres"][dead]) and not np.any(a["contrib"][dead])


def test_oracle_benchmark_unchanged_without_culled_sites():
    case = _heavy(seed=2)
    sp = case.spheres()
    diff = orc.power_face_encode(sp, case.adjacency, case.adjacency_offsets)
    live = orc.live_sites(case.cam.eye.numpy(), sp, case.adjacency, case.adjacency_offsets, diff, kind=1)
    assert (live == 0).mean() > 0.3
    inp = orc.Inputs(sp, case.nsigmas(), case.texel_sites, case.rgbh(), case.adjacency, case.adjacency_offsets, diff)
    cam, full, culled = _lists(case, live)
    a = orc.raster_rgb(cam, inp, *full, threshold=1e-2)
    b = orc.raster_rgb(cam, inp, *culled, threshold=1e-2)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("form", ["exact", "radius", "power"])
def test_culled_spheres_lie_outside_their_cells(form):
    """float64 brute force: every point sampled in (and on) a culled sphere
    lies strictly on the far side of the culling face's plane, as the
    compositing kernels read that plane (exact: the radical plane of the two
    spheres, i.e. the point is nearer in power distance to the neighbour;
    radius / power: the planes of the two binary16 encodings). No point of
    the sphere is then in the site's cell."""
    case = _heavy(seed=3)
    sp32 = case.spheres()
    sp = sp32.astype(np.float64)
    adj, offs = case.adjacency, case.adjacency_offsets
    diff = {"exact": None, "radius": orc.adj_encode(sp32, adj, offs),
            "power": orc.power_face_encode(sp32, adj, offs)}[form]
    live, val, src = orc.live_sites(case.cam.eye.numpy(), sp32, adj, offs, diff, kind=int(form == "power"),
                                    return_margin=True)
    c = sp[src, :3]
    r = sp[src, 3]
    if diff is None:
        nv = sp[adj, :3] - c
        h = 0.5 * ((nv * nv).sum(1) + r * r - sp[adj, 3] ** 2)
    else:
        d = diff.view(np.float16).astype(np.float64)
        nv = d[:, :3]
        h = d[:, 3] - (c * nv).sum(1) if form == "power" else 0.5 * ((nv * nv).sum(1) + r * r - d[:, 3] ** 2)
    rng = np.random.default_rng(0)
    culled_edges = np.nonzero(val > 0)[0]
    assert culled_edges.size > 100
    for e in culled_edges[:400]:
        u = rng.normal(size=(256, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        y = u * r[e] * np.concatenate([np.ones(128), rng.uniform(0, 1, 128) ** (1 / 3)])[:, None]
        assert np.all(y @ nv[e] > h[e]), e
        if diff is None:
            x = y + c[e]
            j = adj[e]
            pi = (y ** 2).sum(1) - r[e] ** 2
            pj = ((x - sp[j, :3]) ** 2).sum(1) - sp[j, 3] ** 2
            assert np.all(pj < pi), e
    # a site is culled exactly when one of its faces is
    dead = np.zeros(case.n, bool)
    dead[src[culled_edges]] = True
    assert np.array_equal(live == 0, dead)


def test_filter_lists_keeps_order():
    off = np.array([0, 3, 3, 6], np.int64)
    sites = np.array([5, 1, 2, 1, 7, 5], np.int32)
    live = np.ones(8, np.uint8)
    live[[1, 7]] = 0
    o2, s2 = orc.filter_lists(off, sites, live)
    assert o2.tolist() == [0, 2, 2, 3]
    assert s2.tolist() == [5, 2, 5]


def _bounded(case, inp):
    eye = case.cam.eye.numpy()
    live = orc.live_sites(eye, case.spheres(), case.adjacency, case.adjacency_offsets, inp.adjacency_diff)
    bounds, empty = orc.site_bounds(eye, case.spheres(), case.adjacency, case.adjacency_offsets,
                                    inp.adjacency_diff, live=live)
    return (~empty).astype(np.uint8), bounds


@pytest.mark.parametrize("exact_geometry", [False, True])
def test_oracle_outputs_unchanged_with_bound_balls(exact_geometry):
    """Lists restricted to the tiles that meet each site's ball (and without
    the sites whose ball vanished): the oracle's compositing forward and
    backward are unchanged, per pixel bit for bit."""
    case = _heavy(seed=4)
    inp = oracle_inputs(case, exact_geometry)
    live, bounds = _bounded(case, inp)
    cam, full, _ = _lists(case, live)
    c = case.cam
    restricted = orc.restrict_lists(cam, c.eye.numpy(), c.right.numpy(), c.up.numpy(), *full, bounds, live)
    assert restricted[1].size < 0.6 * full[1].size
    g = random_grads(case, seed=2)
    res = []
    for off, sites in (full, restricted):
        out = orc.raster_fwd(cam, inp, off, sites, depth_quantiles=case.depth_quantiles, ray_gt=case.gt,
                             return_point_err=True, precision="f32")
        glt = -g["opacity"].astype(np.float64) * np.exp(out["log_t"])
        gr = orc.raster_bwd(cam, inp, off, sites, out["tile_early_stop"], out["log_t"], g["color"], glt,
                            g["normal_distance"], g["normal"], g["contrib"], depth_quantiles=case.depth_quantiles,
                            grad_quantile_depths=g["quantile_depths"], precision="f32")
        res.append((out, gr))
    (a, ga), (b, gb) = res
    for k in ("color", "log_t", "normal_distance", "normal", "quantile_depths"):
        assert np.array_equal(a[k], b[k]), k
    for x, y, k in [(a[k], b[k], k) for k in ("contrib", "point_err")] + [(ga[k], gb[k], k) for k in GRAD_KEYS]:
        np.testing.assert_allclose(y, x, rtol=4e-6, atol=4e-6 * float(np.abs(x).max()), err_msg=k)


@pytest.mark.parametrize("form", ["exact", "radius"])
def test_bound_balls_hold_the_cells(form):
    """float64 brute force: points sampled uniformly in each sphere that lie on
    the cell side of every face (as the kernels read the faces) are inside
    the site's ball; a site whose ball vanished has no such point; balls of
    live sites never exceed the inflated sphere."""
    case = _heavy(seed=5)
    sp32 = case.spheres()
    sp = sp32.astype(np.float64)
    adj, offs = case.adjacency, case.adjacency_offsets
    diff = None if form == "exact" else orc.adj_encode(sp32, adj, offs)
    eye = case.cam.eye.numpy()
    live = orc.live_sites(eye, sp32, adj, offs, diff)
    bounds, empty = orc.site_bounds(eye, sp32, adj, offs, diff, live=live)
    assert (empty & (live == 1)).sum() > 0          # the sweeps find more empty cells
    shrunk = (~empty) & (bounds[:, 3] < 0.9 * sp32[:, 3])
    assert shrunk.sum() > 50
    rng = np.random.default_rng(1)
    checked = 0
    for i in np.nonzero(live)[0][:600]:
        a0, a1 = offs[i], offs[i + 1]
        if diff is None:
            nv = sp[adj[a0:a1], :3] - sp[i, :3]
            h = 0.5 * ((nv * nv).sum(1) + sp[i, 3] ** 2 - sp[adj[a0:a1], 3] ** 2)
        else:
            d = diff[a0:a1].view(np.float16).astype(np.float64)
            nv = d[:, :3]
            h = 0.5 * ((nv * nv).sum(1) + sp[i, 3] ** 2 - d[:, 3] ** 2)
        u = rng.normal(size=(4000, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        y = u * sp[i, 3] * rng.uniform(0, 1, (4000, 1)) ** (1 / 3)
        inside = np.all(y @ nv.T <= h, axis=1) if nv.size else np.ones(4000, bool)
        if empty[i]:
            assert not inside.any(), i
            continue
        p = y[inside] + sp[i, :3]
        dist = np.linalg.norm(p - bounds[i, :3].astype(np.float64), axis=1)
        assert np.all(dist <= bounds[i, 3] * (1 + 1e-6)), i
        checked += int(inside.any())
    assert checked > 100
