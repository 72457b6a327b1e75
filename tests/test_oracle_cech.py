"""The Cech oracle (oracle/cech.py) pinned before it checks the GPU builder:
known answers of the predicate, the kd-tree candidate search against all
pairs, and every disagreement with the exact FP64 predicate a near-tie.
A mistake planted in the oracle's predicate (<= for <, a dropped z term, a
difference of radii) fails these checks."""

import numpy as np
import pytest

from oracle import cech as oc
from powerfoam_amd import synth


def _pairs(adj, off):
    i = np.repeat(np.arange(off.size - 1), np.diff(off))
    return set(zip(i.tolist(), adj.tolist()))


def test_known_answers():
    P = np.array([[0, 0, 0], [1, 0, 0],            # touching: |d| = 1 = 0.5 + 0.5
                  [5, 0, 0], [5, 0, 0],            # coincident centres, r > 0
                  [9, 9, 9], [9, 9, 9],            # coincident centres, r = 0
                  [20, 0, 0], [20.8, 0.8, 0.8],    # boxes overlap, balls do not (|d| = 1.386 > 1)
                  [30, 0, 0], [30.6, 0, 0.8]],     # |d| = 1 < 1.0 + eps: overlap
                 np.float32)
    R = np.array([0.5, 0.5, 0.1, 0.2, 0, 0, 0.5, 0.5, 0.5, 0.50001], np.float32)
    adj, off = oc.cech_adjacency(P, R, brute=True)
    assert _pairs(adj, off) == {(2, 3), (3, 2), (8, 9), (9, 8)}
    assert np.array_equal(off, [0, 0, 0, 1, 2, 2, 2, 2, 2, 3, 4])
    assert oc.cech_adjacency(P, R)[0].tolist() == adj.tolist()


def _check_fp64(P, R, adj, off):
    """Every ordered pair the float32 lists and the exact predicate disagree on
    is within the rounding bound; returns (#pairs, #disagreements)."""
    n = P.shape[0]
    got = _pairs(adj, off)
    ie, je = synth.cech_adjacency(P, R)  # FP64 predicate on the float32 inputs
    want = _pairs(ie, je)
    diff = np.array(sorted(got ^ want), np.int64).reshape(-1, 2)
    if diff.size:
        m, bound = oc.fp64_margin(P, R, diff[:, 0], diff[:, 1])
        assert np.all(np.abs(m) <= bound), (diff[np.abs(m) > bound][:5], m[np.abs(m) > bound][:5])
    assert got == {(j, i) for i, j in got}
    assert all(i != j for i, j in got)
    assert off[-1] == len(got) and off.size == n + 1
    return len(got), len(diff)


def _scene(seed, n):
    s = synth.scene_uniform(seed, n=n)
    return s["points"], s["radii"]


def test_kdtree_candidates_equal_all_pairs():
    rng = np.random.default_rng(3)
    P = rng.uniform(-1, 1, (1500, 3)).astype(np.float32)
    R = np.exp(rng.uniform(np.log(0.005), np.log(0.2), 1500)).astype(np.float32)
    a, o = oc.cech_adjacency(P, R)
    b, p = oc.cech_adjacency(P, R, brute=True)
    assert np.array_equal(a, b) and np.array_equal(o, p)
    assert a.size > 2000


@pytest.mark.parametrize("seed,n", [(0, 50_000), (5, 20_000)])
def test_fp64_predicate_up_to_ties(seed, n):
    P, R = _scene(seed, n)
    adj, off = oc.cech_adjacency(P, R)
    pairs, ties = _check_fp64(P, R, adj, off)
    print(pairs, "pairs,", ties, "near-tie disagreements")
    assert pairs > 10_000


def test_lattice_of_touching_balls():
    """Integer lattice, r = 0.5: every axis neighbour touches exactly (not
    adjacent), and r = 0.5 + 2^-10 makes them adjacent; diagonals never."""
    g = np.stack(np.meshgrid(*[np.arange(6.0)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    adj, off = oc.cech_adjacency(g, np.full(len(g), 0.5, np.float32))
    assert adj.size == 0
    adj, off = oc.cech_adjacency(g, np.full(len(g), 0.5 + 2 ** -10, np.float32))
    assert adj.size == 2 * 3 * 5 * 36  # axis-neighbour pairs of a 6^3 lattice, both ways
    _check_fp64(g, np.full(len(g), 0.5 + 2 ** -10, np.float32), adj, off)


@pytest.mark.parametrize("mistake", ["le", "no_z", "r_diff"])
def test_planted_mistakes_are_caught(mistake, monkeypatch):
    F = np.float32
    good = oc.overlap

    def bad(b, i, j):
        lo, hi, c, r = b
        d = (c[i] - c[j]).astype(F)
        sq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(F)
        if mistake != "no_z":
            sq = (sq + d[:, 2] * d[:, 2]).astype(F)
        s = (r[i] - r[j]) if mistake == "r_diff" else (r[i] + r[j])
        ok = (np.sqrt(sq) <= s) if mistake == "le" else (np.sqrt(sq) < s)
        return ok & (i != j)

    monkeypatch.setattr(oc, "overlap", bad)
    caught = False
    try:
        test_known_answers()
    except AssertionError:
        caught = True
    if not caught:
        P, R = _scene(0, 5000)
        try:
            _check_fp64(P, R, *oc.cech_adjacency(P, R))
        except AssertionError:
            caught = True
    monkeypatch.setattr(oc, "overlap", good)
    assert caught, mistake
