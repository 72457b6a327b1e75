"""Pins of the learned colour head's oracle (oracle/mlp.py), independent of its
formulas:

1. bf16 rounding equals torch's float32 -> bfloat16 conversion (nearest-even);
2. the forward (rounding off) equals a torch float64 network assembled from
   nn.functional.linear/relu on the same encoding;
3. the backward (rounding off) equals torch autograd of that network and FP64
   central differences of the oracle's own forward;
4. the encoding contracts with SH coefficients exactly as the SH head does
   (e(d) . sh = the SH head's pre-clamp colour).
"""

import numpy as np
import pytest
import torch

from oracle import mlp as om
from powerfoam_amd import synth


def _params(seed, n=12):
    rng = np.random.default_rng(seed)
    return dict(latents=rng.normal(0, 0.5, (n, 16)), w1=rng.normal(0, 0.25, (64, 32)), b1=rng.normal(0, 0.1, 64),
                w2=rng.normal(0, 0.18, (64, 64)), b2=rng.normal(0, 0.1, 64), w3=rng.normal(0, 0.2, (3, 64)),
                b3=rng.normal(0.3, 0.1, 3))


def _rows(seed, n=12, r=150):
    rng = np.random.default_rng(seed + 1)
    d = rng.normal(size=(r, 3))
    return rng.integers(0, n, r), d / np.linalg.norm(d, axis=1, keepdims=True)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 20000), rng.normal(0, 1e-3, 20000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0, 65504.0])])
    x = x.astype(np.float32)
    want = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(om.bf16(x), want)


def _torch_net(p, sites, dirs):
    t = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    e = torch.tensor(om.encode(dirs))
    x = torch.cat([t["latents"][torch.tensor(sites)], e], 1)
    h1 = torch.relu(torch.nn.functional.linear(x, t["w1"], t["b1"]))
    h2 = torch.relu(torch.nn.functional.linear(h1, t["w2"], t["b2"]))
    return torch.relu(torch.nn.functional.linear(h2, t["w3"], t["b3"])), t


def test_forward_and_backward_match_torch_autograd():
    p = _params(1)
    sites, dirs = _rows(1)
    out, t = _torch_net(p, sites, dirs)
    rgb, _ = om.forward(p, sites, dirs, rounding=False)
    np.testing.assert_allclose(rgb, out.detach().numpy(), rtol=0, atol=1e-12)
    assert (rgb > 0).mean() > 0.5
    g = np.random.default_rng(2).normal(size=rgb.shape)
    (out * torch.tensor(g)).sum().backward()
    mine = om.backward(p, sites, dirs, g, rounding=False)
    for k in p:
        np.testing.assert_allclose(mine[k], t[k].grad.numpy(), rtol=1e-10, atol=1e-12, err_msg=k)


def test_backward_matches_central_differences():
    p = _params(3)
    sites, dirs = _rows(3)
    g = np.random.default_rng(4).normal(size=(sites.shape[0], 3))
    mine = om.backward(p, sites, dirs, g, rounding=False)
    rng = np.random.default_rng(5)
    tested = 0
    for trial in range(120):
        k = list(p)[trial % len(p)]
        idx = tuple(rng.integers(0, s) for s in p[k].shape)
        base = p[k][idx]
        vals = []
        for eps in (1e-6, -1e-6):
            p[k][idx] = base + eps
            vals.append((om.forward(p, sites, dirs, rounding=False)[0] * g).sum())
        p[k][idx] = base
        fd = (vals[0] - vals[1]) / 2e-6
        assert abs(fd - mine[k][idx]) <= 1e-6 * (1 + abs(fd)), (k, idx, fd, mine[k][idx])
        tested += 1
    assert tested == 120


def test_encoding_contracts_like_the_sh_head():
    rng = np.random.default_rng(6)
    sh = rng.normal(0, 0.3, (50, 16, 3))
    sh[:, 0] += 0.6
    d = rng.normal(size=(50, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    lin = np.einsum("nk,nkc->nc", om.encode(d), sh)
    np.testing.assert_allclose(np.maximum(lin, 0), synth.sh_eval(sh, d), atol=1e-12)


def test_frequency_encoding_known_values():
    d = np.array([[1.0, 0.0, 0.0], [0.0, 0.5, -0.5]])
    e = om.encode(d, "freq")
    assert e.shape == (2, 16)
    want0 = [1, 1, 0, 0, 0, 0, 0, -1, 1, 1, 0, 0, 0, 1, 1, 1]
    want1 = [1, 0, 0.5, -0.5, 0, 1, -1, 1, 0, 0, 0, 0, 0, 1, -1, -1]
    np.testing.assert_allclose(e[0], want0, atol=1e-12)
    np.testing.assert_allclose(e[1], want1, atol=1e-12)
