"""The geometry oracle (oracle/geometry.py) pinned: Morton interleave against
a bit-by-bit loop, codes of the box corners and of a lattice, the order
against a lexicographic sort of the interleaved bits; the interpenetration
regulariser against the closed form for two balls, torch FP64 autograd of
the same sum and central differences, with a planted sign error caught."""

import numpy as np
import pytest

from oracle import geometry as og
from oracle import cech as oc
from powerfoam_amd import synth


def _naive_interleave(x, y, z):
    code = 0
    for b in range(21):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def test_expand_bits_against_bit_loop():
    rng = np.random.default_rng(0)
    v = np.concatenate([[0, 1, 2 ** 21 - 1], rng.integers(0, 2 ** 21, 2000)]).astype(np.uint64)
    got = og.expand_bits(v)
    want = [sum(((int(x) >> b) & 1) << (3 * b) for b in range(21)) for x in v]
    assert [int(g) for g in got] == want


def test_codes_of_corners_and_lattice():
    g = np.stack(np.meshgrid(*[np.arange(8.0)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32) / 7
    codes = og.morton_codes(g)
    assert codes[0] == 0 and codes[-1] == 2 ** 63 - 1
    q = np.minimum(np.floor(np.float32(g) * np.float32(2 ** 21)), 2 ** 21 - 1).astype(np.int64)
    want = [_naive_interleave(*map(int, row)) for row in q]
    assert codes.tolist() == want
    # a flat box: the y extent is floored at 1e-6
    flat = np.array([[0, 5, 0], [1, 5, 1], [0.5, 5, 0.25]], np.float32)
    assert og.morton_codes(flat)[1] == _naive_interleave(2 ** 21 - 1, 0, 2 ** 21 - 1)


def test_order_is_lexicographic_in_interleaved_bits():
    rng = np.random.default_rng(1)
    P = rng.normal(size=(3000, 3)).astype(np.float32)
    P[100:200] = P[:100]  # ties keep index order
    codes = og.morton_codes(P)
    bits = [(codes >> b) & 1 for b in range(63)]  # lexsort: last key is primary
    want = np.lexsort([np.arange(len(P))] + bits)
    assert np.array_equal(og.morton_sort(P), want)


def test_interpenetration_two_balls():
    P = np.array([[0, 0, 0], [0.3, 0.4, 0]], np.float64)
    R = np.array([0.4, 0.35])
    adj, off = np.array([1, 0]), np.array([0, 1, 2])
    a = og.interpenetration(P, R, adj, off)
    assert np.allclose(a, [(0.75 - 0.5) ** 2] * 2)
    gp, gr = og.interpenetration_bwd(P, R, adj, off, np.array([1.0, 0.0]))
    u = (P[0] - P[1]) / 0.5
    assert np.allclose(gp[0], -2 * 0.25 * u) and np.allclose(gp[1], 2 * 0.25 * u)
    assert np.allclose(gr, [0.5, 0.5])
    # coincident centres: no direction term
    gp, gr = og.interpenetration_bwd(np.zeros((2, 3)), R, adj, off, np.array([1.0, 1.0]))
    assert not gp.any() and np.allclose(gr, [4 * 0.75, 4 * 0.75])


def _scene():
    s = synth.scene_uniform(3, n=2000, r_range=(0.03, 0.15))
    adj, off = oc.cech_adjacency(s["points"], s["radii"])
    return s["points"].astype(np.float64), s["radii"].astype(np.float64), adj, off


def test_adjoint_against_torch_autograd_and_differences():
    import torch
    P, R, adj, off = _scene()
    g = np.random.default_rng(2).normal(size=P.shape[0])
    gp, gr = og.interpenetration_bwd(P, R, adj, off, g)
    tp = torch.tensor(P, requires_grad=True)
    tr = torch.tensor(R, requires_grad=True)
    i = torch.tensor(np.repeat(np.arange(P.shape[0]), np.diff(off)))
    j = torch.tensor(adj.astype(np.int64))
    d = (tp[i] - tp[j]).norm(dim=1)
    s = tr[i] + tr[j]
    t = torch.where(d < s, (s - d) ** 2, torch.zeros_like(d))
    (t * torch.tensor(g)[i]).sum().backward()
    assert np.allclose(gp, tp.grad.numpy(), atol=1e-12) and np.allclose(gr, tr.grad.numpy(), atol=1e-12)
    rng = np.random.default_rng(3)
    eps = 1e-6
    for k in rng.choice(P.shape[0], 40, replace=False):
        for c in range(4):
            Pp, Pm, Rp, Rm = P.copy(), P.copy(), R.copy(), R.copy()
            if c < 3:
                Pp[k, c] += eps
                Pm[k, c] -= eps
            else:
                Rp[k] += eps
                Rm[k] -= eps
            fd = (og.interpenetration(Pp, Rp, adj, off) @ g - og.interpenetration(Pm, Rm, adj, off) @ g) / (2 * eps)
            an = gp[k, c] if c < 3 else gr[k]
            assert abs(fd - an) <= 1e-6 * (1 + abs(an)), (k, c, fd, an)


def test_planted_sign_error_is_caught(monkeypatch):
    import torch
    P, R, adj, off = _scene()
    good = og.interpenetration_bwd

    def bad(*a):
        gp, gr = good(*a)
        return -gp, gr

    monkeypatch.setattr(og, "interpenetration_bwd", bad)
    with pytest.raises(AssertionError):
        test_adjoint_against_torch_autograd_and_differences()
