"""pf_prune_faces on the GPU against its CPU restatement (oracle/prune.py),
bit for bit, with the fp16 face encoding and with full-precision spheres;
renders with the pruned faces against renders with every face.

The restatement is pinned in tests/test_oracle_prune.py (FP64 intervals of
random rays unchanged). Renders: the cells inside the balls are the same, so
outputs agree to float rounding -- colour, log T and normals within 1e-6
abs, gradients within 1e-5 relative per tensor (a ray whose FP32 interval
end is decided by a dropped face's rounding is the only way they differ)."""

import numpy as np
import pytest

from oracle import orc
from oracle import prune as op
from powerfoam_amd import synth
from tests import harness

pytestmark = pytest.mark.gpu


def _gpu_prune(sp, adj, off, diff, margin=1e-3):
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd._abi import check, ptr, stream_handle
    lib = _abi.load()
    t = lambda a, dt: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    n, m = sp.shape[0], adj.size
    spt, adjt, offt = t(sp, torch.float32), t(adj, torch.int32), t(off, torch.int32)
    difft = t(diff.view(np.int16), torch.int16) if diff is not None else None
    a_out = torch.empty(m, dtype=torch.int32, device="cuda")
    d_out = torch.empty((m, 4), dtype=torch.int16, device="cuda") if diff is not None else None
    cnt = torch.empty(n, dtype=torch.int32, device="cuda")
    check(lib.pf_prune_faces(ptr(spt), n, ptr(adjt), ptr(offt), m, ptr(difft), float(margin), ptr(a_out), ptr(d_out),
                             ptr(cnt), stream_handle()))
    torch.cuda.synchronize()
    return (a_out.cpu().numpy(), None if d_out is None else d_out.cpu().numpy().view(np.uint16),
            cnt.cpu().numpy())


@pytest.mark.parametrize("encoded", [True, False])
@pytest.mark.parametrize("kind", ["gmm", "dense"])
def test_prune_bitexact_against_restatement(kind, encoded):
    s = synth.scene_gmm(4, n=3000) if kind == "gmm" else synth.scene_uniform(5, n=3000, r_range=(0.05, 0.2))
    adj, off = synth.power_adjacency(s["points"], s["radii"])
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1).astype(np.float32)
    diff = orc.adj_encode(sp, adj, off) if encoded else None
    print(f"prune {kind} encoded={encoded}: {adj.size} faces", flush=True)
    got = _gpu_prune(sp, adj, off, diff)
    want = op.prune(sp, adj, off, diff)
    print("kept", int(got[2].sum()), "of", adj.size, flush=True)
    assert np.array_equal(got[2], want[2])
    assert np.array_equal(got[0], want[0])
    if encoded:
        assert np.array_equal(got[1], want[1])
    assert got[2].sum() < adj.size


@pytest.mark.parametrize("case", [dict(seed=3, n=1500, width=97, height=61, height_std=0.02),
                                  dict(seed=6, n=1500, width=64, height=64, fisheye_fov=180.0)])
def test_renders_with_pruned_faces_match_every_face(case):
    c = harness.small_case(**case)
    gin = harness.random_grads(c, seed=2)
    print(f"render pruned vs every face: {case}", flush=True)
    (o1, g1), (o0, g0) = (harness.run_gpu(c, "cuda", gin, prune_faces=p) for p in (True, False))
    for k in ("color", "opacity", "normal", "normal_distance"):
        err = np.abs(o1[k] - o0[k]).max()
        print(k, err, flush=True)
        assert err < 1e-6, k
    errs = {k: harness.rel_err(g1[k], g0[k]) for k in harness.GRAD_KEYS}
    print(errs, flush=True)
    assert max(errs.values()) < 1e-5, errs
