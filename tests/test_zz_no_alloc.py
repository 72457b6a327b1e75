"""pf_bin_sort_static: the per-tile sort with its long-tile list in the
library's own device array (no workspace argument, nothing allocated by the
library), against pf_bin_sort with a workspace, through the C-ABI. Two calls
in a row on one stream, the second with fewer tiles and fewer long tiles, so
a stale list from the first call would show. Size rejection is checked on the
host, before any device work."""

import numpy as np
import pytest

from powerfoam_amd import _abi


def test_static_sort_rejects_too_many_tiles():
    lib = _abi.load()
    n = _abi.SORT_STATIC_MAX_TILES + 1
    # any non-null pointer: the size check runs before the pointers are used
    st = lib.pf_bin_sort_static(1, n, None, None, None)
    assert st == _abi.PF_ERR_OVERFLOW
    assert b"pf_bin_sort_static" in lib.pf_last_error()
    assert lib.pf_bin_sort_static(None, 4, None, None, None) == _abi.PF_ERR_ARG


def _segments(rng, cnt):
    off = np.zeros(cnt.size + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    m = int(off[-1])
    keys = rng.integers(0, 3000, m).astype(np.uint64)  # many equal depths
    site = rng.permutation(m).astype(np.uint64)
    return off, (keys << np.uint64(32)) | site


@pytest.mark.gpu
def test_static_sort_matches_workspace_sort_twice():
    import torch
    from powerfoam_amd._abi import check, ptr
    lib = _abi.load()
    rng = np.random.default_rng(21)
    big = np.array([0, 5, 4097, 9000, 64, 16384, 20000, 300, 4096, 12000, 1, 7000], np.int64)
    small = np.array([3, 4500, 2, 4100, 0], np.int64)
    for call, cnt in enumerate((big, small)):
        off, rec = _segments(rng, cnt)
        nt, m = cnt.size, int(off[-1])
        print(f"static sort call {call}: {nt} tiles, {m} records, {(cnt > 4096).sum()} long", flush=True)
        want = np.concatenate([np.sort(rec[off[t]:off[t + 1]]) for t in range(nt)]) & np.uint64(0xFFFFFFFF)
        offsets = torch.tensor(off, device="cuda")
        got = {}
        for mode in ("workspace", "static"):
            records = torch.tensor(rec.view(np.int64), device="cuda")
            sites = torch.full((m,), -1, dtype=torch.int32, device="cuda")
            if mode == "workspace":
                wsb = int(lib.pf_bin_sort_workspace_size(m, nt))
                ws = torch.empty(max(wsb // 4, 1), dtype=torch.int32, device="cuda")
                check(lib.pf_bin_sort(ptr(offsets), nt, ptr(records), ptr(sites), ptr(ws), wsb, None))
            else:
                check(lib.pf_bin_sort_static(ptr(offsets), nt, ptr(records), ptr(sites), None))
            torch.cuda.synchronize()
            got[mode] = sites.cpu().numpy().astype(np.int64)
        assert np.array_equal(got["static"], got["workspace"])
        assert np.array_equal(got["static"], want.astype(np.int64))
