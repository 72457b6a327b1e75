"""GPU parity of the compositing kernels against the FP64 oracle, through the
drop-in Rasterizer (forward and autograd backward)."""

import numpy as np
import pytest

from tests import harness

pytestmark = pytest.mark.gpu

ABS_TOL = 1e-4      # RGB, opacity, depth, normals: max abs error vs the FP64 oracle
GRAD_REL_TOL = 1e-3  # gradients: ||gpu - oracle|| / ||oracle|| per tensor, over the
#                      sites the oracle does not flag as ill-conditioned (rays
#                      tangent to the sphere or intervals at the rounding floor,
#                      where one-ulp differences flip whole contributions);
#                      at most MAX_FRAGILE of the sites may be flagged
MAX_FRAGILE = 0.02


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    return torch.device("cuda:0")


CASES = [
    dict(seed=3, n=400, width=48, height=40),
    dict(seed=5, n=1500, width=97, height=61, fov=70.0),   # ragged: out-of-bounds lanes
    dict(seed=7, n=800, width=64, height=48, height_std=0.05, T=5),
    dict(seed=9, n=600, width=40, height=33, quantiles=[0.9, 0.5, 0.1]),
    # ray-map cameras (equidistant fisheye), binned through the cone tree
    dict(seed=11, n=1500, width=56, height=56, eye=(0.0, 0.0, 1.3), fisheye_fov=150.0),
    dict(seed=12, n=1500, width=64, height=64, eye=(0.2, 0.1, 0.6), fisheye_fov=180.0, height_std=0.05),
    # a fifth of the sites sit inside a larger neighbour: the tile lists leave
    # them out (pf_live_sites), the oracle keeps them, as the reference does
    dict(seed=13, n=2000, width=64, height=48, r_range=(0.02, 0.2), quantiles=[0.5, 0.1]),
]


@pytest.mark.parametrize("cfg", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_forward_matches_fp64_oracle(dev, cfg):
    case = harness.small_case(**cfg)
    print(f"forward case {cfg}", flush=True)
    out, _ = harness.run_gpu(case, dev)
    ref, _ = harness.run_oracle(case)
    for k in ("color", "opacity", "normal_distance", "normal"):
        err = np.abs(out[k] - ref[k]).max()
        assert err < ABS_TOL, (k, err)
    if case.depth_quantiles is not None:
        # unfilled entries stay -1 in both
        assert np.array_equal(out["quantile_depths"] == -1, ref["quantile_depths"] == -1)
        err = np.abs(out["quantile_depths"] - ref["quantile_depths"]).max()
        assert err < ABS_TOL, err
    for k in ("contrib", "point_err"):
        err = np.abs(out[k] - ref[k]).max()
        assert err < ABS_TOL * max(1.0, np.abs(ref[k]).max()), (k, err)
    # invariant: sum of per-site weights = opacity summed over pixels / (H W)
    assert abs(out["contrib"].sum() - out["opacity"].mean()) < 1e-4


@pytest.mark.parametrize("cfg", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_backward_matches_fp64_oracle(dev, cfg):
    case = harness.small_case(**cfg)
    print(f"backward case {cfg}", flush=True)
    gin = harness.random_grads(case, seed=cfg["seed"])
    _, grads = harness.run_gpu(case, dev, gin)
    _, ref = harness.run_oracle(case, gin)
    keep = harness.well_conditioned(ref)
    for k in harness.GRAD_KEYS:
        err = harness.rel_err(grads[k][keep], ref[k][keep])
        assert err < GRAD_REL_TOL, (k, err)


@pytest.mark.parametrize("outputs", [("color",), ("opacity", "normal"), ("normal_distance", "contrib")])
def test_backward_with_unused_outputs(dev, outputs):
    """Outputs the loss leaves out reach the backward as None (gradients are
    not materialised): the gradients equal the oracle's with zero incoming
    gradients for those outputs."""
    cfg = CASES[3]
    case = harness.small_case(**cfg)
    print(f"backward from {outputs} only", flush=True)
    gin = harness.random_grads(case, seed=17)
    for k in ("color", "opacity", "normal_distance", "normal", "contrib", "quantile_depths"):
        if k not in outputs:
            gin[k] = np.zeros_like(gin[k])
    _, grads = harness.run_gpu(case, dev, gin, loss_outputs=outputs)
    _, ref = harness.run_oracle(case, gin)
    keep = harness.well_conditioned(ref)
    for k in harness.GRAD_KEYS:
        if not np.any(ref[k][keep]):
            assert not np.any(grads[k][keep]), k
            continue
        err = harness.rel_err(grads[k][keep], ref[k][keep])
        assert err < GRAD_REL_TOL, (k, err)


@pytest.mark.parametrize("cfg,attr", [(CASES[0], "float"), (CASES[1], "float"), (CASES[2], "half"),
                                      (CASES[4], "float"), (CASES[6], "float")],
                         ids=["case0", "case1", "case2-half", "case4-fisheye", "case6-culled"])
def test_benchmark_matches_fp64_oracle(dev, cfg, attr):
    """Rasterizer.benchmark (colour only, per-ray stop, fp16 power faces)."""
    case = harness.small_case(**cfg)
    print(f"benchmark case {cfg} {attr}", flush=True)
    col, diff = harness.run_gpu_benchmark(case, dev, attr_dtype=attr)
    ref = harness.run_oracle_benchmark(case, diff, attr_dtype=attr)
    err = np.abs(col - ref).max()
    assert err < ABS_TOL, err
    assert ref.max() > 0.1


@pytest.mark.parametrize("cfg,opts", [
    (CASES[0], dict(transmittance_threshold=1e-3, max_intersections=1024, depth_quantile=0.5,
                    bkgd_color=(0.1, 0.7, 0.3))),
    (CASES[1], dict(transmittance_threshold=1e-2, max_intersections=7, depth_quantile=0.8, bkgd_color=(1, 1, 1))),
    (CASES[5], dict(transmittance_threshold=1e-3, max_intersections=1024, depth_quantile=0.9,
                    bkgd_color=(0, 0, 0))),
], ids=["case0", "case1-capped", "case5-fisheye"])
def test_visualize_matches_fp64_oracle(dev, cfg, opts):
    """Rasterizer.visualize (visualization_kernel): colour over background,
    depth, normals, alpha within 1e-4; entry counts exact."""
    import torch
    from powerfoam_amd.camera import TorchCamera
    from powerfoam_amd.rasterize import Rasterizer, VisOptions
    from oracle import orc
    case = harness.small_case(**cfg)
    print(f"visualize case {cfg} {opts}", flush=True)
    T = case.texel_sites.shape[1]
    r = Rasterizer(harness._args(T, case.cam.ray_maps is None), dev)
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    cam = TorchCamera(case.cam.eye, case.cam.right, case.cam.up, case.width, case.height,
                      None if case.cam.ray_maps is None else case.cam.ray_maps.to(dev))
    out = r.visualize(cam, tt(case.points), tt(case.radii), tt(case.density), tt(case.normals),
                      tt(case.texel_sites), tt(case.texel_rgb), tt(case.texel_height),
                      tt(case.adjacency, torch.int32), tt(case.adjacency_offsets, torch.int32), VisOptions(**opts))
    color, depth, normal, alpha, inter = [t.cpu().numpy() for t in out]
    ocam, _k = case.pf_camera()
    offsets, sites, _ = orc.bin_lists(ocam, case.spheres())
    ref = orc.raster_vis(ocam, harness.oracle_inputs(case), offsets, sites, threshold=opts["transmittance_threshold"],
                         max_intersections=opts["max_intersections"], depth_quantile=opts["depth_quantile"],
                         bkgd_color=opts["bkgd_color"])
    for k, v in (("color", color), ("depth", depth), ("normal", normal), ("alpha", alpha)):
        err = np.abs(v - ref[k]).max()
        assert err < ABS_TOL, (k, err)
    assert np.array_equal(inter, ref["intersections"])
    assert (ref["depth"] != 0).mean() > 0.01


SPLIT_CASES = [
    dict(seed=3, n=400, width=48, height=40),
    # dense, zoomed in: lists of 100-180 sites, tiles stop inside later segments
    dict(seed=21, n=3000, width=40, height=32, r_range=(0.05, 0.2), fov=25.0, density_scale=5.0),
    dict(seed=22, n=4000, width=48, height=40, r_range=(0.04, 0.15), quantiles=[0.9, 0.5, 0.1], fov=25.0,
         density_scale=5.0, height_std=0.05),
    dict(seed=12, n=1500, width=64, height=64, eye=(0.2, 0.1, 0.6), fisheye_fov=180.0, height_std=0.05),
    # a fifth of the sites sit inside a larger neighbour: the tile lists leave
    # them out (pf_live_sites), the oracle keeps them, as the reference does
    dict(seed=13, n=2000, width=64, height=48, r_range=(0.02, 0.2), quantiles=[0.5, 0.1]),
]


@pytest.mark.parametrize("cfg", SPLIT_CASES, ids=[f"split{i}" for i in range(len(SPLIT_CASES))])
@pytest.mark.parametrize("threshold", [1e-3, 0.3])
@pytest.mark.parametrize("split_backward", [False, True], ids=["bwd-whole", "bwd-split"])
def test_split_lists_match_unsplit_and_oracle(dev, cfg, threshold, split_backward):
    """Lists split into 64-site segments: the forward, the early-stop indices
    and the gradients equal the unsplit walk up to summation order and the
    FP64 oracle within the parity bars; threshold 0.3 stops tiles inside
    segments."""
    case = harness.small_case(**cfg)
    print(f"split case {cfg} thr={threshold}", flush=True)
    gin = harness.random_grads(case, seed=cfg["seed"])
    out_s, g_s = harness.run_gpu(case, dev, gin, threshold=threshold, seg_len=64, split_backward=split_backward)
    out_u, g_u = harness.run_gpu(case, dev, gin, threshold=threshold, seg_len=None)
    ref, gref = harness.run_oracle(case, gin, threshold=threshold)
    for k in ("color", "opacity", "normal_distance", "normal"):
        assert np.abs(out_s[k] - out_u[k]).max() < 2e-6, k
        assert np.abs(out_s[k] - ref[k]).max() < ABS_TOL, k
    if case.depth_quantiles is not None:
        assert np.array_equal(out_s["quantile_depths"] == -1, ref["quantile_depths"] == -1)
        assert np.abs(out_s["quantile_depths"] - ref["quantile_depths"]).max() < ABS_TOL
    assert np.abs(out_s["contrib"] - out_u["contrib"]).max() < 1e-6
    keep = harness.well_conditioned(gref, max_fraction=0.05)  # dense views graze more sites
    for k in harness.GRAD_KEYS:
        assert harness.rel_err(g_s[k], g_u[k]) < 1e-5, k
        assert harness.rel_err(g_s[k][keep], gref[k][keep]) < GRAD_REL_TOL, k


@pytest.mark.parametrize("n,T", [(0, 8), (1, 1), (1000, 8), (300_001, 5)])
def test_pack_sites_bit_exact(dev, n, T):
    """pf_pack_sites equals the reference's concatenations (torch.cat) bit for
    bit, for contiguous and strided inputs."""
    import torch
    from powerfoam_amd.rasterize import Rasterizer
    print(f"pack sites n={n} T={T}", flush=True)
    r = Rasterizer(harness._args(T), dev)
    g = torch.Generator(device="cpu").manual_seed(n + T)
    mk = lambda *shape: torch.randn(*shape, generator=g).to(dev)
    pts, rad, nrm, den = mk(n, 3), mk(n), mk(n, 3), mk(n)
    rgb, h = mk(n, T, 3), mk(T, n).t()  # the heights as a strided view
    sp, ns, rgbh = r.pack_sites(pts, rad, nrm, den, rgb, h)
    torch.cuda.synchronize()
    assert torch.equal(sp, torch.cat([pts, rad[:, None]], -1))
    assert torch.equal(ns, torch.cat([nrm, den[:, None]], -1))
    assert torch.equal(rgbh, torch.cat([rgb, h[..., None]], -1))
    with pytest.raises(ValueError):
        r.pack_sites(pts, rad, nrm, den, rgb, h[:, :0])
