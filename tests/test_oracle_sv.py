"""Pins of the oracle's Spherical-Voronoi colour head (oracle/sv.py,
color_fn.py:54-184):

- closed forms: one lobe gives val + 0.5 (clamped at 0); identical lobes
  give their common value; texels outside the cone of the cut-off are grey
  with weight sum 0; the cut-off angle is the frustum diagonal plus 10
  degrees;
- the exact adjoint equals torch autograd of the forward and FP64 central
  differences; the reference's 1e-3 skip differs from it only on skipped
  lobes; a planted sign error fails the autograd check.
"""

import math

import numpy as np
import pytest

from oracle import sv


def _setup(seed, P=400, D=8):
    rng = np.random.default_rng(seed)
    eye = np.array([0.2, -0.1, 3.0])
    right, up = np.array([0.6, 0.0, 0.0]), np.array([0.0, 0.45, 0.0])
    fwd = sv.camera_forward(right, up)
    pts = rng.uniform(-1, 1, (P, 3))
    ax = rng.normal(size=(D, P, 3))
    temp = np.linalg.norm(ax, axis=-1) * 2.0
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    val = rng.normal(0, 0.3, (D, P, 3))
    return pts, eye, fwd, sv.fov_cos_cutoff(right, up), ax, val, temp


def test_closed_forms():
    pts, eye, fwd, cut, ax, val, temp = _setup(0, D=1)
    rgb, ws = sv.forward(pts, eye, fwd, -1.0, ax, val, temp)
    np.testing.assert_allclose(rgb, np.maximum(val[0] + 0.5, 0), atol=1e-14)
    pts, eye, fwd, cut, ax, val, temp = _setup(1, D=4)
    val[:] = val[0]
    rgb, _ = sv.forward(pts, eye, fwd, -1.0, ax, val, temp)
    np.testing.assert_allclose(rgb, np.maximum(val[0] + 0.5, 0), atol=1e-12)
    # points behind the camera are outside the cone: grey, weight sum 0
    behind = eye[None] + np.array([[0.0, 0.0, 1.0], [0.5, 0.5, 2.0]])
    rgb, ws = sv.forward(behind, eye, fwd, sv.fov_cos_cutoff([0.6, 0, 0], [0, 0.45, 0]), ax[:, :2], val[:, :2],
                         temp[:, :2])
    np.testing.assert_array_equal(rgb, 0.5)
    np.testing.assert_array_equal(ws, 0.0)
    assert abs(sv.fov_cos_cutoff([1.0, 0, 0], [0, 0, 0]) - math.cos(math.radians(55.0))) < 1e-12


def _torch_forward(pts, eye, fwd, cut, ax, val, temp):
    import torch
    d = torch.tensor(pts) - torch.tensor(eye)[None]
    d = d / d.norm(dim=-1, keepdim=True)
    dist = (d[None] - ax).norm(dim=-1)
    w = torch.exp(-temp * dist)
    rgb = (w[..., None] * val).sum(0) / w.sum(0)[:, None] + 0.5
    return torch.clamp(rgb, min=0.0)


def test_exact_adjoint_matches_autograd_and_differences():
    import torch
    pts, eye, fwd, cut, ax, val, temp = _setup(2)
    g = np.random.default_rng(3).normal(size=(pts.shape[0], 3))
    t = [torch.tensor(a, requires_grad=True) for a in (ax, val, temp)]
    out = _torch_forward(pts, eye, fwd, cut, *t)
    (out * torch.tensor(g)).sum().backward()
    rgb, ws = sv.forward(pts, eye, fwd, -1.0, ax, val, temp)
    np.testing.assert_allclose(rgb, out.detach().numpy(), atol=1e-13)
    ga, gv, gt = sv.backward(pts, eye, fwd, -1.0, ax, val, temp, rgb, ws, g, exact=True)
    np.testing.assert_allclose(ga, t[0].grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gv, t[1].grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gt, t[2].grad.numpy(), rtol=1e-9, atol=1e-12)
    # central differences on a few lobe parameters
    loss = lambda a_, v_, t_: (sv.forward(pts, eye, fwd, -1.0, a_, v_, t_)[0] * g).sum()
    rng = np.random.default_rng(4)
    eps = 1e-6
    for arr, grad in ((ax, ga), (val, gv), (temp, gt)):
        flat, gf = arr.reshape(-1), grad.reshape(-1)
        for j in rng.choice(flat.size, 10, replace=False):
            old = flat[j]
            flat[j] = old + eps
            lp = loss(ax, val, temp)
            flat[j] = old - eps
            lm = loss(ax, val, temp)
            flat[j] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(gf[j] - fd) < 1e-6 * (1 + abs(fd)), (j, gf[j], fd)
    # the reference's skip of lobes below 1e-3 of the weight sum
    ga2, gv2, gt2 = sv.backward(pts, eye, fwd, -1.0, ax, val, temp, rgb, ws, g)
    e = pts[None] - eye[None, None]
    e /= np.linalg.norm(e, axis=-1, keepdims=True)
    w = np.exp(-temp * np.linalg.norm(e - ax, axis=-1))
    skipped = w / ws[None] < 1e-3
    assert skipped.any() and not skipped.all()
    np.testing.assert_allclose(gv2[~skipped], gv[~skipped], rtol=1e-12)
    assert np.all(gv2[skipped] == 0) and np.all(gt2[skipped] == 0)


def test_planted_sign_error_is_caught(monkeypatch):
    """A direction of the wrong sign in the adjoint (d - a read as -d - a)
    fails the autograd check."""
    import torch
    pts, eye, fwd, cut, ax, val, temp = _setup(5, P=50)
    g = np.random.default_rng(6).normal(size=(pts.shape[0], 3))
    t = [torch.tensor(a, requires_grad=True) for a in (ax, val, temp)]
    (_torch_forward(pts, eye, fwd, cut, *t) * torch.tensor(g)).sum().backward()
    rgb, ws = sv.forward(pts, eye, fwd, -1.0, ax, val, temp)
    real = sv._dirs
    monkeypatch.setattr(sv, "_dirs", lambda p, e: -real(p, e))
    ga, gv, gt = sv.backward(pts, eye, -fwd, -1.0, ax, val, temp, rgb, ws, g, exact=True)
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(ga, t[0].grad.numpy(), rtol=1e-9, atol=1e-12)


def test_layout_matches_scene_att_sv():
    """att_from_params restates get_att_sv (scene.py:290-312): lobe k of texel
    (n, t) sits at [k, n*T + t]."""
    rng = np.random.default_rng(7)
    N, T, D = 5, 3, 4
    raw = rng.normal(size=(N, T, 3 * D))
    rgb = rng.normal(size=(N, T, 3 * D))
    ax, val, temp = sv.att_from_params(raw, rgb, D)
    for n in range(N):
        for t in range(T):
            for k in range(D):
                r = raw[n, t, 3 * k:3 * k + 3]
                np.testing.assert_allclose(ax[k, n * T + t], r / np.linalg.norm(r))
                np.testing.assert_allclose(temp[k, n * T + t], np.linalg.norm(r))
                np.testing.assert_allclose(val[k, n * T + t], rgb[n, t, 3 * k:3 * k + 3])
