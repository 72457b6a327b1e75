"""Failure surfacing on the GPU, forced from the host with inputs the
kernels' own guards handle (every kernel exits normally):

- an emit pass given offsets that disagree with its count (one tile's
  segment two records short, the reference's `idx < offsets[tile+1]` guard,
  rasterize.py:358) writes every other record, drops the two, counts them
  in the `dropped` word, and the Rasterizer raises PfError at its next sync;
- a list whose edge count disagrees with adjacency_offsets[-1] is refused;
- NaN and Inf site parameters: every gradient the backward returns is finite
  (the non-finite scrub, rasterize.py:1954-1958), and sites that share no
  tile with a poisoned site, nor neighbour a site that does, get the
  gradients of the clean scene.
"""

import numpy as np
import pytest

from tests import harness

pytestmark = pytest.mark.gpu


def _rast(T=8):
    from powerfoam_amd.rasterize import Rasterizer
    return Rasterizer(harness._args(T), "cuda")


def test_emit_with_short_segment_counts_dropped_records():
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd._abi import check, ptr, stream_handle
    from powerfoam_amd.rasterize import _ref
    case = harness.small_case(seed=4, n=600)
    r = _rast()
    lib = r.lib
    cam, _ = r._camera(case.cam)
    sp = torch.tensor(case.spheres(), device="cuda")
    n = sp.shape[0]
    th, tw = _abi.tiles_hw(case.height, case.width)
    nt = th * tw
    counts = torch.empty(nt + 1, dtype=torch.int32, device="cuda")
    check(lib.pf_bin_count(_ref(cam), ptr(sp), None, None, n, ptr(counts), stream_handle()))
    offsets = torch.empty(nt + 1, dtype=torch.int64, device="cuda")
    check(lib.pf_bin_scan(ptr(counts), nt, ptr(offsets), stream_handle()))
    off = offsets.cpu().numpy()
    cnt = np.diff(off)
    t_short = int(np.argmax(cnt))
    assert cnt[t_short] >= 3
    print(f"emit with tile {t_short} two records short of its {cnt[t_short]}", flush=True)
    off_short = off.copy()
    off_short[t_short + 1:] -= 2
    results = {}
    for name, o in (("exact", off), ("short", off_short)):
        d_off = torch.tensor(o, device="cuda")
        m = int(o[-1])
        records = torch.zeros(m, dtype=torch.int64, device="cuda")
        cursor = torch.empty(nt, dtype=torch.int32, device="cuda")
        dropped = torch.zeros(1, dtype=torch.int32, device="cuda")
        check(lib.pf_bin_emit(_ref(cam), ptr(sp), None, None, n, ptr(d_off), nt, ptr(records), ptr(cursor), None,
                              ptr(dropped), stream_handle()))
        sites = torch.empty(m, dtype=torch.int32, device="cuda")
        check(lib.pf_bin_sort_static(ptr(d_off), nt, ptr(records), ptr(sites), stream_handle()))
        torch.cuda.synchronize()
        results[name] = (o, sites.cpu().numpy(), int(dropped.item()))
    o0, s0, d0 = results["exact"]
    o1, s1, d1 = results["short"]
    assert d0 == 0 and d1 == 2
    for t in range(nt):
        a = set(s0[o0[t]:o0[t + 1]].tolist())
        b = set(s1[o1[t]:o1[t + 1]].tolist())
        if t == t_short:
            assert b <= a and len(a - b) == 2
        else:
            assert a == b
    # the Rasterizer reports a dropped count at its next sync
    r._dropped = torch.full((1,), 2, dtype=torch.int32, device="cuda")
    with pytest.raises(_abi.PfError, match="dropped 2"):
        r.check_binning()
    r.check_binning()  # reported once, then cleared


def test_inconsistent_adjacency_is_refused():
    import torch
    case = harness.small_case(seed=2, n=300)
    r = _rast()
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    adj = tt(case.adjacency, torch.int32)
    off = case.adjacency_offsets.copy()
    with pytest.raises(ValueError, match="adjacency_offsets"):
        r.forward(case.cam, None, tt(case.points), tt(case.radii), tt(case.density), tt(case.normals),
                  tt(case.texel_sites), tt(case.texel_rgb), tt(case.texel_height), adj[:-1], tt(off, torch.int32),
                  tt(case.gt), False)
    torch.cuda.synchronize()


def test_nonfinite_parameters_give_finite_gradients():
    import torch
    case = harness.small_case(seed=8, n=500)
    gin = harness.random_grads(case, seed=3)
    clean_out, clean = harness.run_gpu(case, "cuda", gin)
    # poison the site that appears in the most tiles, and another one's texel colour
    r = _rast()
    cam = case.cam
    sp = torch.tensor(case.spheres(), device="cuda")
    offsets, sites, _, nt = r.bin(cam, sp)
    off, st = offsets.cpu().numpy(), sites.cpu().numpy()
    freq = np.bincount(st, minlength=case.n)
    k, j = (int(v) for v in np.argsort(freq)[-2:])
    print(f"non-finite parameters: density of site {k} NaN, texel rgb of site {j} +Inf", flush=True)
    bad = harness.Case(**{f: getattr(case, f).copy() if isinstance(getattr(case, f), np.ndarray) else getattr(case, f)
                          for f in harness.Case.__dataclass_fields__})
    bad.density[k] = np.nan
    bad.texel_rgb[j, 0, :] = np.inf
    out, grads = harness.run_gpu(bad, "cuda", gin)
    for key in harness.GRAD_KEYS:
        assert np.all(np.isfinite(grads[key])), key
    # sites that share no tile with k or j are untouched by the poison
    tile_of = np.repeat(np.arange(nt), np.diff(off))
    hit = np.unique(tile_of[(st == k) | (st == j)])
    touched = np.zeros(case.n, bool)
    touched[st[np.isin(tile_of, hit)]] = True
    # face gradients also reach the neighbours of the sites in those tiles
    src = np.repeat(np.arange(case.n), np.diff(case.adjacency_offsets))
    touched[case.adjacency[touched[src]]] = True
    free = ~touched
    assert free.sum() > 50
    for key in harness.GRAD_KEYS:
        a, b = grads[key][free], clean[key][free]
        assert harness.rel_err(a, b) < 1e-5, key
