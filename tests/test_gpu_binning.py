"""GPU tile binning is bit-exact against the FP32 oracle mirror: the same
per-tile offsets, the same depth-sorted site lists, the same visibility mask."""

import numpy as np
import pytest

from oracle import orc
from powerfoam_amd import synth
from powerfoam_amd.camera import look_at

pytestmark = pytest.mark.gpu


def _bin_gpu(cam_t, spheres):
    import torch
    from types import SimpleNamespace
    from powerfoam_amd.camera import TorchCamera
    from powerfoam_amd.rasterize import Rasterizer
    pin = cam_t.ray_maps is None
    r = Rasterizer(SimpleNamespace(num_texel_sites=8, render_objective="volume", is_pinhole=pin), "cuda")
    sp = torch.tensor(spheres, device="cuda")
    cam = TorchCamera(cam_t.eye, cam_t.right, cam_t.up, cam_t.width, cam_t.height,
                      None if pin else cam_t.ray_maps.cuda())
    offsets, sites, visible, nt = r.bin(cam, sp)
    return offsets.cpu().numpy(), sites.cpu().numpy(), visible.cpu().numpy()


def _check(cam_t, spheres):
    rm = None if cam_t.ray_maps is None else cam_t.ray_maps.numpy()
    cam, _ = orc.camera(cam_t.eye.numpy(), cam_t.right.numpy(), cam_t.up.numpy(), cam_t.width, cam_t.height, rm)
    off_o, sites_o, _ = orc.bin_lists(cam, spheres)
    off_g, sites_g, vis_g = _bin_gpu(cam_t, spheres)
    assert np.array_equal(off_g, off_o)
    assert np.array_equal(sites_g, sites_o)
    vis_o = np.zeros(spheres.shape[0], bool)
    vis_o[sites_o] = True
    assert np.array_equal(vis_g, vis_o)
    return off_o[-1]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_binning_bitexact_uniform(seed):
    s = synth.scene_uniform(seed, n=20000, r_range=(0.005, 0.2), extent=3.0)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    cam = look_at((0.3, 0.2, 3.0), (0.0, 0.0, 0.0), 331, 217, 65.0)
    m = _check(cam, sp)
    assert m > 0


def test_binning_bitexact_gmm_view():
    s = synth.scene_gmm(1, n=100_000)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    cam = synth.circle_cameras(n_views=32, width=1297, img_h=840)[5]
    m = _check(cam, sp)
    assert m > 100_000


def test_binning_empty_and_offscreen():
    # every site behind the camera or culled: all tiles empty
    cam = look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 64, 48, 60.0)
    sp = np.array([[0.0, 0.0, 5.0, 0.1], [0.0, 0.0, 2.9, 0.5], [50.0, 0.0, 0.0, 0.1]], np.float32)
    assert _check(cam, sp) == 0


def _fisheye_cam(eye, size, fov):
    import torch
    from powerfoam_amd.camera import fisheye_ray_maps
    cam = look_at(eye, (0.0, 0.0, 0.0), size, size, 60.0)
    cam.ray_maps = torch.tensor(fisheye_ray_maps(eye, (0.0, 0.0, 0.0), size, fov))
    return cam


@pytest.mark.parametrize("size,fov", [(37, 180.0), (256, 120.0), (1024, 180.0)])
def test_cone_tree_bitexact(size, fov):
    import torch
    from powerfoam_amd import _abi
    cam_t = _fisheye_cam((0.3, 0.2, 2.5), size, fov)
    rm = cam_t.ray_maps.numpy()
    cam, _k = orc.camera(cam_t.eye.numpy(), cam_t.right.numpy(), cam_t.up.numpy(), size, size, rm)
    ref = orc.cones(cam)
    lib = _abi.load()
    assert lib.pf_cone_tree_nodes(size, size) == ref.shape[0]
    rmd = cam_t.ray_maps.cuda().contiguous()
    from powerfoam_amd.camera import TorchCamera
    pc = TorchCamera(cam_t.eye, cam_t.right, cam_t.up, size, size, rmd).to_pf()
    out = torch.empty(ref.shape, dtype=torch.float32, device="cuda")
    _abi.check(lib.pf_bin_cones(_abi.ctypes.byref(pc), out.data_ptr(), _abi.stream_handle()))
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("seed,size,fov", [(0, 96, 180.0), (1, 333, 120.0)])
def test_cone_binning_bitexact(seed, size, fov):
    s = synth.scene_uniform(seed, n=20000, r_range=(0.005, 0.2), extent=3.0)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    m = _check(_fisheye_cam((0.3, 0.2, 3.0), size, fov), sp)
    assert m > 1000


def test_cone_binning_bitexact_gmm_fisheye():
    s = synth.scene_gmm(1, n=100_000)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    m = _check(_fisheye_cam((4.0, 1.0, 0.0), 1024, 180.0), sp)
    assert m > 100_000


@pytest.mark.parametrize("seg_len", [64, 512, 1024])
def test_tile_schedule_orders_and_segment_table(seg_len):
    """pf_tile_schedule: tile order and segment order are permutations sorted by
    descending length bucket (exact below 512, steps of 32 up to 16,864, one
    bucket above); the segment table is the exclusive scan of
    max(1, ceil(count / seg_len)) with each segment naming its tile."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd._abi import check, ptr
    lib = _abi.load()
    rng = np.random.default_rng(seg_len)
    nt = 17115
    cnt = np.minimum(rng.lognormal(4.0, 1.6, nt).astype(np.int64), 40000)
    cnt[rng.random(nt) < 0.05] = 0
    off = np.zeros(nt + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    m = int(off[-1])
    cap = nt + m // seg_len
    dev = torch.device("cuda")
    i32 = dict(dtype=torch.int32, device=dev)
    offsets = torch.tensor(off, device=dev)
    order, first = torch.empty(nt, **i32), torch.empty(nt + 1, **i32)
    seg_tile, seg_order, count = torch.empty(cap, **i32), torch.empty(cap, **i32), torch.empty(1, **i32)
    print(f"tile schedule nt={nt} m={m} seg_len={seg_len}", flush=True)
    check(lib.pf_tile_schedule(ptr(offsets), nt, m, seg_len, cap, ptr(order), ptr(first), ptr(seg_tile),
                               ptr(seg_order), ptr(count), None))
    torch.cuda.synchronize()

    def bucket(n):
        n = np.asarray(n, np.int64)
        return np.where(n < 512, n, np.minimum(512 + ((n - 512) >> 5), 1023))

    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(nt))
    assert np.all(np.diff(bucket(cnt[o])) <= 0)
    ns = np.maximum(1, (cnt + seg_len - 1) // seg_len)
    want_first = np.concatenate([[0], np.cumsum(ns)])
    assert np.array_equal(first.cpu().numpy(), want_first)
    total = int(want_first[-1])
    assert int(count.item()) == total <= cap
    st = seg_tile.cpu().numpy()[:total]
    assert np.array_equal(st, np.repeat(np.arange(nt), ns))
    so = seg_order.cpu().numpy()[:total]
    assert np.array_equal(np.sort(so), np.arange(total))
    size = cnt[st] - (np.arange(total) - want_first[st]) * seg_len
    size = np.minimum(size, seg_len)
    assert np.all(np.diff(np.minimum(size[so], 1023)) <= 0)


@pytest.mark.parametrize("nt", [1, 1000, 17115, 70000])
def test_tile_schedule_order_only(nt):
    """pf_tile_schedule without the segment outputs (the order the compositing
    kernels take when no list is split): a permutation of the tiles sorted by
    descending length bucket, also for tile counts that are not a multiple of
    the 1024 threads and above 64 tiles per thread."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd._abi import check, ptr
    lib = _abi.load()
    rng = np.random.default_rng(nt)
    cnt = np.minimum(rng.lognormal(4.0, 1.6, nt).astype(np.int64), 40000)
    cnt[rng.random(nt) < 0.05] = 0
    off = np.zeros(nt + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    offsets = torch.tensor(off, device="cuda")
    order = torch.full((nt,), -1, dtype=torch.int32, device="cuda")
    print(f"tile order nt={nt}", flush=True)
    check(lib.pf_tile_schedule(ptr(offsets), nt, int(off[-1]), 0, 0, ptr(order), None, None, None, None, None))
    torch.cuda.synchronize()
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(nt))
    b = np.where(cnt < 512, cnt, np.minimum(512 + ((cnt - 512) >> 5), 1023))
    assert np.all(np.diff(b[o]) <= 0)


def test_sort_long_tiles_with_and_without_workspace():
    """pf_bin_sort on segments from 1 to 40,000 records (the 16k-record LDS path
    and its global-merge tail above that): with the workspace (long tiles in
    the second kernel) and without it, the sites equal numpy's sort of the
    64-bit records, duplicate keys ordered by site id."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd._abi import check, ptr
    lib = _abi.load()
    rng = np.random.default_rng(11)
    cnt = np.array([0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 5000, 9000, 16384, 16385, 23000, 40000, 7, 300],
                   np.int64)
    off = np.zeros(cnt.size + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    m = int(off[-1])
    keys = rng.integers(0, 2000, m).astype(np.uint64)  # many ties
    site = rng.permutation(m).astype(np.uint64)
    rec = (keys << np.uint64(32)) | site
    want = np.concatenate([np.sort(rec[off[t]:off[t + 1]]) for t in range(cnt.size)]) & np.uint64(0xFFFFFFFF)
    nt = cnt.size
    for use_ws in (True, False):
        print(f"long-tile sort workspace={use_ws}", flush=True)
        offsets = torch.tensor(off, device="cuda")
        records = torch.tensor(rec.view(np.int64), device="cuda")
        sites = torch.empty(m, dtype=torch.int32, device="cuda")
        wsb = int(lib.pf_bin_sort_workspace_size(m, nt)) if use_ws else 0
        ws = torch.empty(max(wsb // 4, 1), dtype=torch.int32, device="cuda")
        check(lib.pf_bin_sort(ptr(offsets), nt, ptr(records), ptr(sites), ptr(ws) if use_ws else None, wsb, None))
        torch.cuda.synchronize()
        got = sites.cpu().numpy().astype(np.int64)
        assert np.array_equal(got, want.astype(np.int64))


@pytest.mark.parametrize("n", [1, 300, 5000])
def test_adjacency_encoding_bit_exact(n):
    """pf_adj_encode (prefetch_adjacency_kernel, rasterize.py:580-606) against the
    oracle's binary16 encoding, bit for bit: degrees from 0 to 150 (edges of one
    site spread over several passes of a wave), site counts not a multiple of
    the 256 sites a workgroup takes."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd._abi import check, ptr
    lib = _abi.load()
    rng = np.random.default_rng(n)
    sp = np.concatenate([rng.normal(0, 1, (n, 3)), rng.uniform(0.01, 0.2, (n, 1))], 1).astype(np.float32)
    deg = rng.integers(0, 12, n)
    deg[rng.random(n) < 0.05] = 0
    if n > 10:
        deg[rng.integers(0, n, 3)] = 150
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(deg)
    adj = rng.integers(0, n, int(off[-1])).astype(np.int32)
    want = orc.adj_encode(sp, adj, off.astype(np.int32))
    print(f"adjacency encode n={n} edges={adj.size}", flush=True)
    d_sp = torch.tensor(sp, device="cuda")
    d_adj = torch.tensor(adj, device="cuda")
    d_off = torch.tensor(off.astype(np.int32), device="cuda")
    out = torch.zeros((max(adj.size, 1), 4), dtype=torch.int16, device="cuda")
    check(lib.pf_adj_encode(ptr(d_sp), n, ptr(d_adj), ptr(d_off), adj.size, ptr(out), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint16)[: adj.size]
    assert np.array_equal(got, want)
