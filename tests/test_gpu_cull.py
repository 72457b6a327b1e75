"""GPU side of the site culling: pf_live_sites against the oracle predicate,
and binning with the mask bit-exact against the oracle's lists with the same
sites left out (the visibility mask unchanged). Compositing parity with the
culled lists is covered by the culled case of test_gpu_raster.py."""

import ctypes
import functools

import numpy as np
import pytest

from oracle import orc
from powerfoam_amd import synth
from tests import harness

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=1)
def _scenes():
    c = harness.small_case(13, n=2000, width=64, height=48, r_range=(0.02, 0.2))
    s = synth.scene_gmm(1, n=30_000)
    adj, off = synth.power_adjacency(s["points"], s["radii"])
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1).astype(np.float32)
    return [("cech2k", c.spheres(), c.adjacency, c.adjacency_offsets, c.cam),
            ("gmm30k", sp, adj, off, synth.circle_cameras(n_views=32, width=331, img_h=217)[3])]


def _gpu_live(sp, adj, off, cam_t, form, with_bounds=False):
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd.camera import TorchCamera
    from powerfoam_amd.rasterize import power_face_encoding
    lib = _abi.load()
    d_sp = torch.tensor(sp, device="cuda")
    d_adj = torch.tensor(np.asarray(adj, np.int32), device="cuda")
    d_off = torch.tensor(np.asarray(off, np.int32), device="cuda")
    if form == "radius":
        diff = torch.empty((max(len(adj), 1), 4), dtype=torch.float16, device="cuda")
        _abi.check(lib.pf_adj_encode(_abi.ptr(d_sp), sp.shape[0], _abi.ptr(d_adj), _abi.ptr(d_off), len(adj),
                                     _abi.ptr(diff), None))
    elif form == "power":
        diff = power_face_encoding(d_sp[:, :3], d_sp[:, 3], d_adj, d_off)
    else:
        diff = None
    cam = TorchCamera(cam_t.eye, cam_t.right, cam_t.up, cam_t.width, cam_t.height).to_pf()
    live = torch.full((sp.shape[0],), 7, dtype=torch.uint8, device="cuda")
    bounds = torch.full((sp.shape[0], 4), float("nan"), device="cuda") if with_bounds else None
    kind = _abi.PF_FACES_POWER if form == "power" else _abi.PF_FACES_RADIUS
    _abi.check(lib.pf_live_sites(ctypes.byref(cam), _abi.ptr(d_sp), sp.shape[0], _abi.ptr(d_adj), _abi.ptr(d_off),
                                 len(adj), _abi.ptr(diff), kind, _abi.ptr(live), _abi.ptr(bounds), None))
    torch.cuda.synchronize()
    dh = None if diff is None else diff.cpu().numpy().view(np.uint16)[: len(adj)]
    if with_bounds:
        return live.cpu().numpy(), dh, bounds.cpu().numpy()
    return live.cpu().numpy(), dh


@pytest.mark.parametrize("form", ["exact", "radius", "power"])
def test_live_sites_match_oracle(form):
    for name, sp, adj, off, cam in _scenes():
        print(f"live sites {name} {form}", flush=True)
        live_g, diff = _gpu_live(sp, adj, off, cam, form)
        assert set(np.unique(live_g)) <= {0, 1}
        live_o, val, src = orc.live_sites(cam.eye.numpy(), sp, adj, off, diff, kind=int(form == "power"),
                                          return_margin=True)
        assert (live_o == 0).mean() > 0.15, name
        # the float32 decision can differ only for a face within rounding of
        # the margin (the square roots and the order of the products)
        for s in np.nonzero(live_g != live_o)[0]:
            assert np.min(np.abs(val[src == s])) < 1e-6, (name, s)
        assert np.count_nonzero(live_g != live_o) <= 2, name


def test_binning_with_live_mask_bitexact():
    import torch
    from types import SimpleNamespace
    from powerfoam_amd.camera import TorchCamera
    from powerfoam_amd.rasterize import Rasterizer
    for name, sp, adj, off, cam_t in _scenes():
        print(f"binning with live mask {name}", flush=True)
        r = Rasterizer(SimpleNamespace(num_texel_sites=8, render_objective="volume", is_pinhole=True), "cuda")
        d_sp = torch.tensor(sp, device="cuda")
        d_adj = torch.tensor(np.asarray(adj, np.int32), device="cuda")
        d_off = torch.tensor(np.asarray(off, np.int32), device="cuda")
        cam = TorchCamera(cam_t.eye, cam_t.right, cam_t.up, cam_t.width, cam_t.height)
        diff = r.encode_adjacency(d_sp, d_adj, d_off)
        r.bound_sites = False
        live, bounds = r.live_sites(cam, d_sp, d_adj, d_off, diff)
        assert live is not None and bounds is None
        offsets, sites, visible, nt = r.bin(cam, d_sp, live)
        live_h = live.cpu().numpy()
        cam_o, _ = orc.camera(cam_t.eye.numpy(), cam_t.right.numpy(), cam_t.up.numpy(), cam_t.width, cam_t.height)
        off_o, sites_o, _ = orc.bin_lists(cam_o, sp)
        off_f, sites_f = orc.filter_lists(off_o, sites_o, live_h)
        assert np.array_equal(offsets.cpu().numpy(), off_f)
        assert np.array_equal(sites.cpu().numpy(), sites_f)
        assert sites_f.size < 0.9 * sites_o.size, name
        # the visibility mask is the footprint's, culled sites included
        vis_o = np.zeros(sp.shape[0], bool)
        vis_o[sites_o] = True
        assert np.array_equal(visible.cpu().numpy(), vis_o)


@pytest.mark.parametrize("form", ["exact", "radius", "power"])
def test_site_bounds_match_oracle(form):
    """pf_live_sites with balls: the mask (sites whose ball vanished included)
    equals the oracle's up to faces within rounding of a margin, and each
    live site's ball equals the oracle's to float rounding."""
    for name, sp, adj, off, cam in _scenes():
        print(f"site bounds {name} {form}", flush=True)
        live_g, diff, b_g = _gpu_live(sp, adj, off, cam, form, with_bounds=True)
        live0 = orc.live_sites(cam.eye.numpy(), sp, adj, off, diff, kind=int(form == "power"))
        b_o, empty = orc.site_bounds(cam.eye.numpy(), sp, adj, off, diff, kind=int(form == "power"), live=live0)
        live_o = (~empty).astype(np.uint8)
        assert (live_o == 0).mean() > 0.15, name
        assert np.count_nonzero(live_g != live_o) <= max(2, sp.shape[0] // 5000), name
        both = (live_g == 1) & (live_o == 1)
        scale = sp[both, 3] + 1e-6 * (1 + np.abs(sp[both, :3]).max(1))
        err = np.abs(b_g[both] - b_o[both]).max(1) / scale
        assert np.quantile(err, 0.999) < 1e-4 and err.max() < 1e-2, (name, err.max())


def test_binning_with_bounds_bitexact():
    """Lists binned with the mask and the balls equal the oracle's lists
    restricted by the same mask and balls; the visibility mask is the
    sphere footprint's, as without them."""
    import torch
    from types import SimpleNamespace
    from powerfoam_amd.camera import TorchCamera
    from powerfoam_amd.rasterize import Rasterizer
    for name, sp, adj, off, cam_t in _scenes():
        print(f"binning with bounds {name}", flush=True)
        r = Rasterizer(SimpleNamespace(num_texel_sites=8, render_objective="volume", is_pinhole=True), "cuda")
        d_sp = torch.tensor(sp, device="cuda")
        d_adj = torch.tensor(np.asarray(adj, np.int32), device="cuda")
        d_off = torch.tensor(np.asarray(off, np.int32), device="cuda")
        cam = TorchCamera(cam_t.eye, cam_t.right, cam_t.up, cam_t.width, cam_t.height)
        diff = r.encode_adjacency(d_sp, d_adj, d_off)
        live, bounds = r.live_sites(cam, d_sp, d_adj, d_off, diff)
        assert bounds is not None
        offsets, sites, visible, nt = r.bin(cam, d_sp, live, bounds)
        cam_o, _ = orc.camera(cam_t.eye.numpy(), cam_t.right.numpy(), cam_t.up.numpy(), cam_t.width, cam_t.height)
        off_o, sites_o, _ = orc.bin_lists(cam_o, sp)
        off_r, sites_r = orc.restrict_lists(cam_o, cam_t.eye.numpy(), cam_t.right.numpy(), cam_t.up.numpy(),
                                            off_o, sites_o, bounds.cpu().numpy(), live.cpu().numpy())
        assert np.array_equal(offsets.cpu().numpy(), off_r)
        assert np.array_equal(sites.cpu().numpy(), sites_r)
        vis_o = np.zeros(sp.shape[0], bool)
        vis_o[sites_o] = True
        assert np.array_equal(visible.cpu().numpy(), vis_o)


def test_rasterizer_with_and_without_culling_on_power_scene():
    """On the at-scale scene's geometry (GMM sites, power adjacency with hidden
    sites, 331x217 view): the rasterizer with culled sites, ball binning and
    backward entry skipping gives the outputs and gradients of the one that
    bins and walks everything (up to float summation order), the same
    visibility mask, and shorter lists."""
    import torch
    from types import SimpleNamespace
    from powerfoam_amd.camera import TorchCamera
    from powerfoam_amd.rasterize import Rasterizer
    name, sp, adj, off, cam_t = _scenes()[1]
    s = synth.scene_gmm(1, n=30_000)
    rng = np.random.default_rng(0)
    dev = torch.device("cuda")
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    rgb = rng.uniform(0, 1, (sp.shape[0], 8, 3)).astype(np.float32)
    cam = TorchCamera(cam_t.eye, cam_t.right, cam_t.up, cam_t.width, cam_t.height)
    gt = tt(rng.uniform(0, 1, (cam_t.height, cam_t.width, 3)))
    res = []
    for fast in (False, True):
        r = Rasterizer(SimpleNamespace(num_texel_sites=8, render_objective="volume", is_pinhole=True), dev)
        r.cull_hidden = r.bound_sites = r.skip_unshaded = fast
        params = [tt(s["points"]).requires_grad_(), tt(s["radii"]).requires_grad_(),
                  tt(s["density"]).requires_grad_(), tt(s["normals"]).requires_grad_(),
                  tt(s["texel_sites"]).requires_grad_(), tt(rgb).requires_grad_(),
                  tt(s["texel_height"]).requires_grad_()]
        out = r.forward(cam, None, *params, tt(adj, torch.int32), tt(off, torch.int32), gt, True)
        loss = ((out[0] - gt) ** 2).mean() + out[1].mean() + out[3].abs().mean()
        loss.backward()
        torch.cuda.synchronize()
        res.append(([o.detach().cpu().numpy() if o is not None else None for o in out],
                    [q.grad.cpu().numpy() for q in params], int(r.last_tiles[1].numel())))
    (o0, g0, m0), (o1, g1, m1) = res
    print(f"culling on {name}: {m0} -> {m1} list entries", flush=True)
    assert m1 < 0.9 * m0
    for k in (0, 1, 2, 3):  # colour, opacity, normal distance, normal
        assert np.abs(o0[k] - o1[k]).max() < 1e-5, k
    for k in (6, 7):  # contrib, point error (sums over tiles)
        assert np.abs(o0[k] - o1[k]).max() <= 1e-5 * max(1.0, np.abs(o0[k]).max()), k
    assert np.array_equal(o0[8], o1[8])  # prim_visible_mask
    for a, b in zip(g0, g1):
        assert np.linalg.norm(a - b) <= 1e-5 * max(np.linalg.norm(a), 1e-30)
