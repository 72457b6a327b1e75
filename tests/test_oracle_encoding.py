"""Pins of the oracle's binary16 encoding of neighbour deltas
(prefetch_adjacency_kernel, rasterize.py:580-606) against numpy's IEEE
float16 round-to-nearest-even, and of the SH basis against scipy's complex
spherical harmonics."""

import numpy as np
import pytest

from oracle import orc
from powerfoam_amd import synth


def test_f16_rounding_matches_numpy():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.normal(0, 1, 20000), rng.normal(0, 1e-5, 20000), rng.normal(0, 3e4, 5000),
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 6.1e-5,
         np.inf, -np.inf],
    ]).astype(np.float32)
    # exact ties at every scale
    m = rng.integers(0, 2 ** 10, 4000)
    e = rng.integers(-14, 15, 4000)
    ties = ((1024 + m) * 2.0 ** (e - 10) + 2.0 ** (e - 11)).astype(np.float32)
    x = np.concatenate([x, ties, -ties])
    got = orc.f32_to_f16_bits(x)
    want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    nan = orc.f32_to_f16_bits(np.array([np.nan], np.float32))[0]
    assert (nan & 0x7C00) == 0x7C00 and (nan & 0x3FF) != 0


def test_adjacency_encoding():
    s = synth.scene_uniform(1, n=2000, r_range=(0.02, 0.1))
    adj, off = synth.cech_adjacency(s["points"], s["radii"])
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1).astype(np.float32)
    diff = orc.adj_encode(sp, adj, off)
    src = np.repeat(np.arange(sp.shape[0]), np.diff(off))
    want = np.concatenate([sp[adj, :3] - sp[src, :3], sp[adj, 3:4]], 1).astype(np.float16).view(np.uint16)
    assert np.array_equal(diff, want)


def test_cech_adjacency_is_overlap_graph():
    s = synth.scene_uniform(2, n=800, r_range=(0.02, 0.15))
    adj, off = synth.cech_adjacency(s["points"], s["radii"])
    p = s["points"].astype(np.float64)
    r = s["radii"].astype(np.float64)
    dist = np.linalg.norm(p[:, None] - p[None], axis=-1)
    brute = (dist < r[:, None] + r[None]) & ~np.eye(p.shape[0], dtype=bool)
    for i in range(p.shape[0]):
        nb = adj[off[i]:off[i + 1]]
        assert np.all(np.diff(nb) > 0)
        assert set(nb.tolist()) == set(np.nonzero(brute[i])[0].tolist())


def _real_sh_scipy(l, m, d):
    """Real orthonormal SH from scipy's complex ones (Condon-Shortley phase)."""
    from scipy.special import sph_harm
    theta = np.arctan2(d[:, 1], d[:, 0])  # azimuth
    phi = np.arccos(np.clip(d[:, 2], -1, 1))  # polar
    if m > 0:
        return np.sqrt(2) * (-1) ** m * np.real(sph_harm(m, l, theta, phi))
    if m < 0:
        return np.sqrt(2) * (-1) ** m * np.imag(sph_harm(-m, l, theta, phi))
    return np.real(sph_harm(0, l, theta, phi))


def test_sh_basis_matches_scipy():
    """Each basis function equals (up to the sign convention of the constants)
    the orthonormal real SH Y_l^m of scipy, for l = 1..3."""
    import warnings
    rng = np.random.default_rng(3)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ours = synth.sh_basis(d)
    order = [(1, -1), (1, 0), (1, 1), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2),
             (3, -3), (3, -2), (3, -1), (3, 0), (3, 1), (3, 2), (3, 3)]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for k, (l, m) in enumerate(order):
            ref = _real_sh_scipy(l, m, d)
            sign = np.sign(np.sum(ours[:, k] * ref))
            np.testing.assert_allclose(ours[:, k], sign * ref, atol=1e-12)
