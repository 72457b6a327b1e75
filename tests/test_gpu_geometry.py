"""GPU Morton ordering (pf_morton_sort) and interpenetration regulariser
(pf_interpenetration_fwd / _bwd) against the oracle (oracle/geometry.py):
codes and order bit-exact on the 1M-site GMM scene, a uniform scene with
repeated points, a flat box, one point and none; areas within 1e-5 and
gradients within 1e-4 (relative) of FP64 on the 50k-site uniform scene (Cech
lists) and on the 1M-site GMM scene (its GPU Cech lists)."""

import numpy as np
import pytest

from oracle import cech as oc
from oracle import geometry as og
from powerfoam_amd import synth
from tests import harness

pytestmark = pytest.mark.gpu


def _morton(P):
    import torch
    from powerfoam_amd.geometry import morton_codes_and_order, morton_sort
    codes, order = morton_codes_and_order(torch.tensor(P, device="cuda"))
    assert torch.equal(morton_sort(torch.tensor(P, device="cuda")), order)
    return codes.cpu().numpy(), order.cpu().numpy()


@pytest.mark.parametrize("case", ["gmm1m", "uniform_repeats", "flat", "one", "none"])
def test_morton_bitexact(case):
    rng = np.random.default_rng(5)
    if case == "gmm1m":
        P = synth.scene_gmm(1, n=1_000_000)["points"]
    elif case == "uniform_repeats":
        P = rng.uniform(-3, 7, (200_000, 3)).astype(np.float32)
        P[1000:5000] = P[:4000]
        P[9000:9100] = P[9000]
    elif case == "flat":
        P = rng.uniform(-1, 1, (5000, 3)).astype(np.float32)
        P[:, 2] = 0.5
    elif case == "one":
        P = np.array([[1.0, 2.0, 3.0]], np.float32)
    else:
        P = np.zeros((0, 3), np.float32)
    print(f"morton {case} n={len(P)}", flush=True)
    codes, order = _morton(P)
    if len(P) == 0:
        assert codes.size == 0 and order.size == 0
        return
    assert np.array_equal(codes, og.morton_codes(P))
    assert np.array_equal(order, og.morton_sort(P))


def _interp(P, R, adj, off, g, pairing=0):
    import torch
    from powerfoam_amd.geometry import InterpenetrationFunction
    tp = torch.tensor(P, device="cuda", requires_grad=True)
    tr = torch.tensor(R, device="cuda", requires_grad=True)
    a = InterpenetrationFunction.apply(tp, tr, torch.as_tensor(adj, device="cuda"), torch.as_tensor(off, device="cuda"),
                                       pairing)
    (a * torch.tensor(g, dtype=torch.float32, device="cuda")).sum().backward()
    return a.detach().cpu().numpy(), tp.grad.cpu().numpy(), tr.grad.cpu().numpy()


def _check(P, R, adj, off, name, pairing=0):
    print(f"interpenetration {name}: n={len(P)} edges={np.asarray(adj).size} pairing={pairing}", flush=True)
    g = np.random.default_rng(1).normal(size=len(P)).astype(np.float32)
    a, gp, gr = _interp(P, R, adj, off, g, pairing)
    adj_np, off_np = np.asarray(adj), np.asarray(off)
    want = og.interpenetration(P, R, adj_np, off_np)
    wp, wr = og.interpenetration_bwd(P, R, adj_np, off_np, g)
    assert np.abs(a - want).max() <= 1e-5 * np.abs(want).max()
    assert harness.rel_err(gp, wp) < 1e-4 and harness.rel_err(gr, wr) < 1e-4
    assert (want > 0).mean() > 0.3


@pytest.mark.parametrize("pairing", [0, 1, 2])
def test_interpenetration_uniform_50k(pairing):
    s = synth.scene_uniform(0, n=50_000)
    adj, off = oc.cech_adjacency(s["points"], s["radii"])
    _check(s["points"], s["radii"], adj, off, "uniform 50k", pairing)


def test_interpenetration_one_way_lists_paired_by_search():
    """Restricted lists: ascending, partly one-way (sites outside their
    cells); pairing 1 finds the reverse entries that exist."""
    import torch
    from powerfoam_amd.bvh import restricted_power_adjacency
    from tests.test_oracle_power import _nested_case
    case, _ = _nested_case()
    ra, ro = restricted_power_adjacency(torch.tensor(case.points, device="cuda"),
                                        torch.tensor(case.radii, device="cuda"))
    adj, off = ra.cpu().numpy(), ro.cpu().numpy()
    i = np.repeat(np.arange(case.n), np.diff(off))
    pairs = set(zip(i.tolist(), adj.tolist()))
    assert any((b, a) not in pairs for a, b in pairs)  # some entries are one-way
    _check(case.points, case.radii, adj, off, "restricted", 1)


def test_interpenetration_gmm_1m():
    import torch
    from powerfoam_amd.bvh import AABBTree
    s = synth.scene_gmm(1, n=1_000_000)
    t = AABBTree("cuda")
    t.update(torch.tensor(s["points"], device="cuda"), torch.tensor(s["radii"], device="cuda"))
    adj, off = t.build_cech_complex()
    assert adj._pf_pairing == 2
    _check(s["points"], s["radii"], adj.cpu().numpy(), off.cpu().numpy(), "gmm 1M", 2)


def test_interpenetration_no_edges_and_coincident():
    P = np.array([[0, 0, 0], [0, 0, 0], [5, 5, 5]], np.float32)
    R = np.array([0.5, 0.25, 1.0], np.float32)
    adj, off = np.array([1, 0], np.int32), np.array([0, 1, 2, 2], np.int32)
    a, gp, gr = _interp(P, R, adj, off, np.ones(3, np.float32))
    assert np.allclose(a, [0.5625, 0.5625, 0.0]) and not gp.any() and np.allclose(gr, [3.0, 3.0, 0.0])
