"""GPU restricted power adjacency (pf_power_restrict) against the FP64 oracle
(oracle/orc_power.h): keep flags and per-site counts bit-exact on the
50k-site uniform scene, a 30k-site GMM scene, sites nested inside larger
ones, and lists longer than the kernel's LDS stage (degrees up to ~900); on
the 1M-site GMM scene (its GPU Cech lists), 400 sampled sites' faces equal the
oracle's edge test and the counts match the flags; the GPU rasterizer renders
the restricted lists like the full Cech lists."""

import time

import numpy as np
import pytest

from oracle import cech as oc
from oracle import orc
from powerfoam_amd import synth
from tests import harness

pytestmark = pytest.mark.gpu


def _gpu_restrict(P, R, adj, off):
    import torch
    from powerfoam_amd import _abi
    lib = _abi.load()
    sp = torch.tensor(np.concatenate([P, R[:, None]], 1), dtype=torch.float32, device="cuda")
    a = torch.tensor(adj, dtype=torch.int32, device="cuda")
    o = torch.tensor(off, dtype=torch.int32, device="cuda")
    m = adj.size
    keep = torch.full((max(m, 1),), 7, dtype=torch.uint8, device="cuda")
    counts = torch.full((P.shape[0] + 1,), -1, dtype=torch.int32, device="cuda")
    _abi.check(lib.pf_power_restrict(_abi.ptr(sp), P.shape[0], _abi.ptr(a), _abi.ptr(o), m, _abi.ptr(keep),
                                     _abi.ptr(counts), _abi.stream_handle()))
    return keep[:m].cpu().numpy(), counts.cpu().numpy()


def _equal(P, R, adj, off, name):
    print(f"power restrict {name}: n={P.shape[0]} entries={adj.size} max degree={np.diff(off).max()}", flush=True)
    keep, counts = _gpu_restrict(P, R, adj, off)
    want_keep, want_counts = orc.power_restrict(np.concatenate([P, R[:, None]], 1), adj, off)
    assert np.array_equal(counts, want_counts)
    assert np.array_equal(keep, want_keep)
    return int(keep.sum())


def test_uniform_50k_bitexact():
    s = synth.scene_uniform(0, n=50_000)
    adj, off = oc.cech_adjacency(s["points"], s["radii"])
    assert _equal(s["points"], s["radii"], adj, off, "uniform 50k") > 50_000


def test_gmm_30k_bitexact():
    s = synth.scene_gmm(1, n=30_000)
    adj, off = oc.cech_adjacency(s["points"], s["radii"])
    assert _equal(s["points"], s["radii"], adj, off, "gmm 30k") > 30_000


def test_long_lists_bitexact():
    s = synth.scene_uniform(7, n=5000, r_range=(0.2, 0.4))
    adj, off = oc.cech_adjacency(s["points"], s["radii"])
    assert (np.diff(off) > 512).sum() > 100
    _equal(s["points"], s["radii"], adj, off, "long lists")


def test_nested_sites_and_edge_inputs():
    from tests.test_oracle_power import _nested_case
    case, small = _nested_case()
    _equal(case.points, case.radii, case.adjacency, case.adjacency_offsets, "nested")
    # no edges, one site, empty scene
    P = np.random.default_rng(0).uniform(size=(50, 3)).astype(np.float32) * 100
    R = np.full(50, 0.1, np.float32)
    keep, counts = _gpu_restrict(P, R, np.zeros(0, np.int32), np.zeros(51, np.int32))
    assert keep.size == 0 and not counts.any()
    keep, counts = _gpu_restrict(np.zeros((0, 3), np.float32), np.zeros(0, np.float32), np.zeros(0, np.int32),
                                 np.zeros(1, np.int32))
    assert counts.tolist() == [0]


def test_gmm_1m_sampled():
    import torch
    from powerfoam_amd.bvh import AABBTree, restricted_power_adjacency
    s = synth.scene_gmm(1, n=1_000_000)
    P, R = s["points"], s["radii"]
    pt, rt = torch.tensor(P, device="cuda"), torch.tensor(R, device="cuda")
    tree = AABBTree("cuda")
    tree.update(pt, rt)
    adj, off = tree.build_cech_complex()
    print("power restrict 1M gmm", flush=True)
    restricted_power_adjacency(pt, rt, adj, off)
    torch.cuda.synchronize()
    t0 = time.time()
    radj, roff = restricted_power_adjacency(pt, rt, adj, off)
    torch.cuda.synchronize()
    print(f"restrict {1e3 * (time.time() - t0):.1f} ms: {adj.numel()} Cech entries -> {radj.numel()}", flush=True)
    from powerfoam_amd import _abi
    lib = _abi.load()
    sp = torch.cat([pt, rt[:, None]], 1).contiguous()
    keep = torch.empty(adj.numel(), dtype=torch.uint8, device="cuda")
    counts = torch.empty(P.shape[0] + 1, dtype=torch.int32, device="cuda")
    _abi.check(lib.pf_power_restrict(_abi.ptr(sp), P.shape[0], _abi.ptr(adj), _abi.ptr(off), adj.numel(),
                                     _abi.ptr(keep), _abi.ptr(counts), _abi.stream_handle()))
    keep, counts, a, o = keep.cpu().numpy(), counts.cpu().numpy(), adj.cpu().numpy(), off.cpu().numpy()
    assert np.array_equal(np.cumsum(counts), roff.cpu().numpy())
    c = np.concatenate([[0], np.cumsum(keep, dtype=np.int64)])
    assert np.array_equal(c[o[1:]] - c[o[:-1]], counts[1:])
    sp_np = np.concatenate([P, R[:, None]], 1)
    rng = np.random.default_rng(0)
    checked = 0
    for x in rng.choice(P.shape[0], 400, replace=False):
        nb = a[o[x]:o[x + 1]]
        want = np.array([orc.power_face(sp_np, min(x, b), max(x, b), a[o[min(x, b)]:o[min(x, b) + 1]])
                         for b in nb], bool)
        got = keep[o[x]:o[x + 1]].astype(bool)
        if not want.any() and got.any():
            continue  # a site whose cell misses its ball: one-way faces only
        assert np.array_equal(got, want), x
        checked += nb.size
    assert checked > 5_000


def test_rasterizer_renders_restricted_like_cech():
    import torch
    from powerfoam_amd.bvh import restricted_power_adjacency
    from tests.test_oracle_power import _nested_case
    case, _ = _nested_case()
    outs = []
    for restricted in (False, True):
        c = case
        if restricted:
            ra, ro = restricted_power_adjacency(torch.tensor(case.points, device="cuda"),
                                                torch.tensor(case.radii, device="cuda"),
                                                torch.tensor(case.adjacency, device="cuda"),
                                                torch.tensor(case.adjacency_offsets, device="cuda"))
            c = harness.Case(**{**case.__dict__, "adjacency": ra.cpu().numpy(), "adjacency_offsets": ro.cpu().numpy()})
        print(f"raster restricted={restricted} entries={c.adjacency.size}", flush=True)
        out, _ = harness.run_gpu(c, "cuda", exact_geometry=True)
        outs.append(out)
    for k in ("color", "opacity", "normal"):
        assert np.abs(outs[0][k] - outs[1][k]).max() < 2e-5, k
