"""The compositing's face pruning (oracle/prune.py, the restatement of
pf_prune_faces) changes no ray's interval inside its site's ball and cell.

Pinned by FP64 brute force: for every site of two 3,000-site scenes (the GMM
distribution of the benchmark, sparse at this size, and a dense uniform one,
8.7 faces per site) with the power/alpha adjacency the benchmark uses (fp16
face encoding of pf_adj_encode, decoded to FP64), random rays through the ball -- from
outside, grazing, and from inside -- get the same FP64 interval from every
face and from the kept faces only (|difference| < 1e-9, or both empty). A
planted mistake (the disc radius added instead of subtracted in the witness
test) must fail the same check, and the pruning must leave out a real share
of the faces."""

import numpy as np
import pytest

from oracle import orc
from oracle import prune as op
from powerfoam_amd import synth

N_SITES = 3000


@pytest.fixture(scope="module", params=["gmm", "dense"])
def scene(request):
    if request.param == "gmm":
        s = synth.scene_gmm(4, n=N_SITES)
    else:
        s = synth.scene_uniform(5, n=N_SITES, r_range=(0.05, 0.2))
    adj, off = synth.power_adjacency(s["points"], s["radii"])
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1).astype(np.float32)
    diff = orc.adj_encode(sp, adj, off)
    return sp, adj, off, diff


def _faces64(sp, diff, s, edges):
    """FP64 (n, w) of the given edges from the decoded fp16 encoding."""
    c = sp[s].astype(np.float64)
    out = []
    for e in edges:
        v = diff[e].view(np.float16).astype(np.float64)
        n = v[:3]
        out.append((n, 0.5 * (n @ n + c[3] ** 2 - v[3] ** 2)))
    return out


def _rays(rng, c, r, k):
    """k rays through the ball: from outside towards a point inside, grazing
    (aimed at the rim), and from inside."""
    rays = []
    for i in range(k):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        kind = i % 3
        if kind == 0:
            o = c + 3.0 * r * u
            p = c + r * rng.uniform(0.0, 1.0) ** (1 / 3) * v
        elif kind == 1:
            o = c + 3.0 * r * u
            p = c + 0.999 * r * v
        else:
            o = c + 0.7 * r * u
            p = c + r * v
        d = p - o
        rays.append((o, d / np.linalg.norm(d)))
    return rays


def _check(sp, adj, off, diff, drop_fn, rays_per_site=12, seed=0):
    """Number of (site, ray) pairs whose interval changes when the faces
    drop_fn(s) marks are left out."""
    rng = np.random.default_rng(seed)
    bad = 0
    for s in range(sp.shape[0]):
        a0, a1 = int(off[s]), int(off[s + 1])
        if a1 == a0:
            continue
        d = drop_fn(s)
        if not d.any():
            continue
        c, r = sp[s, :3].astype(np.float64), float(sp[s, 3])
        every = _faces64(sp, diff, s, range(a0, a1))
        kept = [f for f, x in zip(every, d) if not x]
        for o, dr in _rays(rng, c, r, rays_per_site):
            a = op.interval(o, dr, c, r, every)
            b = op.interval(o, dr, c, r, kept)
            if (a is None) != (b is None) or (a is not None and max(abs(a[0] - b[0]), abs(a[1] - b[1])) > 1e-9):
                bad += 1
    return bad


def test_pruned_faces_change_no_interval(scene):
    sp, adj, off, diff = scene
    drops = {}

    def drop(s):
        drops[s] = op.dropped(sp, adj, diff, off, s, 1e-3)
        return drops[s]

    assert _check(sp, adj, off, diff, drop) == 0
    n_drop = sum(int(v.sum()) for v in drops.values())
    frac = n_drop / adj.size
    print(f"dropped {n_drop} of {adj.size} faces ({frac:.1%})")
    assert frac > 0.05


def test_planted_mistake_changes_intervals(scene):
    """The witness test with the disc radius added (a disc tested by its far
    side instead of its near side) drops faces that do bound intervals."""
    sp, adj, off, diff = scene

    def wrong(s):
        a0, a1 = int(off[s]), int(off[s + 1])
        c = sp[s].astype(np.float64)
        fs = _faces64(sp, diff, s, range(a0, a1))
        out = []
        for i, (nf, wf) in enumerate(fs):
            drop = False
            if wf > 0:
                nf2 = nf @ nf
                if wf * wf < c[3] ** 2 * nf2:
                    rho = np.sqrt(c[3] ** 2 - wf * wf / nf2)
                    for k, (ng, wg) in enumerate(fs):
                        if k == i or wg <= 0:
                            continue
                        dd = ng @ nf
                        perp = np.sqrt(max(ng @ ng - dd * dd / nf2, 0.0))
                        if dd * wf / nf2 + rho * perp - wg > 0:  # planted: + rho
                            drop = True
                            break
            out.append(drop)
        return np.array(out, bool)

    assert _check(sp, adj, off, diff, wrong) > 0


def test_restatement_writes_kept_faces_first(scene):
    sp, adj, off, diff = scene
    adj_out, diff_out, count = op.prune(sp[:400], adj[: off[400]], off[:401], diff[: off[400]])
    for s in range(400):
        a0, a1 = int(off[s]), int(off[s + 1])
        d = op.dropped(sp, adj, diff, off, s, 1e-3)
        assert count[s] == int((~d).sum())
        assert list(adj_out[a0:a0 + count[s]]) == list(adj[a0:a1][~d])
        assert sorted(adj_out[a0:a1]) == sorted(adj[a0:a1])
        assert (diff_out[a0:a1] == diff[a0:a1][np.r_[np.nonzero(~d)[0], np.nonzero(d)[0][::-1]]]).all()
