"""The restricted power adjacency oracle (oracle/orc_power.h) pinned before
it checks the GPU kernel:

* every face it finds is a weighted Delaunay edge of the lifted convex hull
  (scipy Qhull, the reference's construction in benchmark.py:25-83), and it
  drops only the hull edges whose face misses the balls;
* an independent brute force of the same 2-D question (every chord of a
  constraint line through the disk clipped by all the other half-planes, and
  the disk centre) gives the same answer on sampled pairs;
* the FP64 oracle rasterizer renders the same image and gradients with the
  restricted lists as with the full Cech lists (faces outside the balls never
  bound a ray's interval), including sites nested inside larger ones;
* known answers: nested balls, a face shadowed by a third site, two balls
  alone.
"""

import numpy as np
import pytest

from oracle import cech as oc
from oracle import orc
from powerfoam_amd import synth
from tests import harness


def _sp(P, R):
    return np.concatenate([P, R[:, None]], 1).astype(np.float32)


def _lp_faces(sp, adj, off):
    faces = set()
    for a in range(off.size - 1):
        nb = adj[off[a]:off[a + 1]]
        for b in nb[nb > a]:
            if orc.power_face(sp, a, b, nb):
                faces.add((a, int(b)))
    return faces


@pytest.mark.parametrize("kind", ["uniform", "gmm"])
def test_faces_are_weighted_delaunay_edges(kind):
    s = synth.scene_uniform(0, n=20_000) if kind == "uniform" else synth.scene_gmm(1, n=30_000)
    P, R = s["points"], s["radii"]
    adj, off = oc.cech_adjacency(P, R)
    faces = _lp_faces(_sp(P, R), adj, off)
    p, r = P.astype(np.float64), R.astype(np.float64)
    e, _low = synth._power_edges(p, r)
    hull = set(map(tuple, np.sort(e, 1).tolist()))
    d = np.linalg.norm(p[e[:, 0]] - p[e[:, 1]], axis=1)
    alpha = set(map(tuple, np.sort(e[d < r[e[:, 0]] + r[e[:, 1]]], 1).tolist()))
    print(kind, len(faces), "faces,", len(alpha), "overlap-filtered hull edges")
    assert faces <= hull
    assert 0.6 * len(alpha) < len(faces) < len(alpha)


def _brute_face(P, R, a, b, nb):
    """Disk of the radical plane clipped by the half-planes of nb (FP64): is
    it non-empty? Non-empty iff the disk centre is feasible or some
    constraint line's chord, clipped by all the other half-planes, is."""
    p, r = P.astype(np.float64), R.astype(np.float64)
    d = p[b] - p[a]
    L = np.linalg.norm(d)
    u = d / L
    s = (L * L + r[a] ** 2 - r[b] ** 2) / (2 * L)
    rho2 = r[a] ** 2 - s * s
    if rho2 <= 0:
        return False, np.inf
    t = np.eye(3)[np.argmin(np.abs(u))]
    e1 = np.cross(u, t)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u, e1)
    ks = np.array([k for k in nb if k != b], np.int64)
    if ks.size == 0:
        return True, np.inf
    dk = p[ks] - p[a]
    A = np.stack([dk @ e1, dk @ e2], 1)
    beta = ((dk * dk).sum(1) + r[a] ** 2 - r[ks] ** 2) / 2 - s * (dk @ u)
    slack0 = beta.min() if ks.size else np.inf
    if slack0 >= 0:
        return True, slack0
    best = -np.inf
    for i in range(ks.size):
        nn = A[i] @ A[i]
        if nn == 0:
            continue
        x = A[i] * beta[i] / nn
        h2 = rho2 - x @ x
        if h2 < 0:
            continue
        h = np.sqrt(h2)
        tv = np.array([-A[i, 1], A[i, 0]]) / np.sqrt(nn)
        den = A @ tv
        num = beta - A @ x
        lo, hi = -h, h
        other = np.arange(ks.size) != i
        pos, neg, zer = other & (den > 0), other & (den < 0), other & (den == 0)
        if pos.any():
            hi = min(hi, (num[pos] / den[pos]).min())
        if neg.any():
            lo = max(lo, (num[neg] / den[neg]).max())
        if zer.any() and num[zer].min() < 0:
            continue
        best = max(best, hi - lo)
    return best >= 0, best


def test_brute_force_agrees_on_sampled_pairs():
    s = synth.scene_gmm(3, n=20_000)
    P, R = s["points"], s["radii"]
    adj, off = oc.cech_adjacency(P, R)
    sp = _sp(P, R)
    rng = np.random.default_rng(0)
    sites = rng.choice(P.shape[0], 5000, replace=False)
    n_pairs = n_kept = 0
    for a in sites:
        nb = adj[off[a]:off[a + 1]]
        for b in nb[nb > a]:
            want, margin = _brute_face(P, R, a, b, nb)
            got = orc.power_face(sp, a, b, nb)
            if got != want:
                # only a chord of length ~0 (a face of zero width) may differ
                assert abs(margin) < 1e-9, (a, b, got, want, margin)
            n_pairs += 1
            n_kept += got
    print(n_pairs, "pairs,", n_kept, "faces")
    assert n_pairs > 2000 and 0.2 * n_pairs < n_kept < n_pairs


def _nested_case():
    case = harness.small_case(seed=4, n=900, width=72, height=56, r_range=(0.03, 0.2))
    # sites nested inside larger ones: their cells miss their balls
    rng = np.random.default_rng(9)
    big = np.argsort(case.radii)[-40:]
    P = case.points.copy()
    R = case.radii.copy()
    small = rng.choice(np.setdiff1d(np.arange(case.n), big), 40, replace=False)
    P[small] = P[big] + rng.normal(0, 0.02, (40, 3)).astype(np.float32)
    R[small] = (0.3 * R[big]).astype(np.float32)
    case.points, case.radii = P, R
    case.adjacency, case.adjacency_offsets = oc.cech_adjacency(P, R)
    return case, small


def test_restricted_lists_render_like_cech_lists():
    case, small = _nested_case()
    keep, counts = orc.power_restrict(case.spheres(), case.adjacency, case.adjacency_offsets)
    grads_in = harness.random_grads(case, seed=2)
    full, gfull = harness.run_oracle(case, grads_in, exact_gradients=True, exact_geometry=True)
    n_full = case.adjacency.size
    case.adjacency = case.adjacency[keep.astype(bool)]
    case.adjacency_offsets = np.cumsum(counts).astype(np.int32)
    print(n_full, "Cech entries ->", case.adjacency.size)
    assert case.adjacency.size < 0.8 * n_full
    # the nested sites kept their one-way faces
    assert all(counts[k + 1] > 0 for k in small)
    red, gred = harness.run_oracle(case, grads_in, exact_gradients=True, exact_geometry=True)
    for k in ("color", "opacity", "normal", "normal_distance"):
        assert np.abs(full[k] - red[k]).max() < 1e-10, k
    for k in harness.GRAD_KEYS:
        assert np.abs(gfull[k] - gred[k]).max() <= 1e-9 * (1 + np.abs(gfull[k]).max()), k


def test_known_answers():
    # 0: big ball; 1: nested inside it; 2, 3: two equal overlapping balls far
    # away; 4, 5, 6: a face shadowed by a third site (6 sits between 4 and 5)
    P = np.array([[0, 0, 0], [0.1, 0, 0], [5, 0, 0], [5.8, 0, 0], [10, 0, 0], [10.9, 0, 0], [10.45, 0, 0]],
                 np.float32)
    R = np.array([1.0, 0.3, 0.5, 0.5, 0.5, 0.5, 0.45], np.float32)
    adj, off = oc.cech_adjacency(P, R)
    keep, counts = orc.power_restrict(_sp(P, R), adj, off)
    pairs = set(zip(np.repeat(np.arange(7), np.diff(off)).tolist(), adj.tolist()))
    kept = {e for e, k in zip(sorted(pairs), keep) if k}
    assert (1, 0) in kept and (0, 1) not in kept            # the nested site keeps a one-way face
    assert {(2, 3), (3, 2)} <= kept
    assert (4, 5) in pairs and (4, 5) not in kept and (5, 4) not in kept
    assert {(4, 6), (6, 4), (5, 6), (6, 5)} <= kept
