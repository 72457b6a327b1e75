"""Pins of the oracle's scene glue (oracle/glue.py, scene.py:239-337):

- the normal / tangent / bitangent are the rows of scipy's rotation matrix of
  the same quaternion (an independent library), orthonormal and right-handed;
- softplus equals torch.nn.functional.softplus(beta=100) on both sides of
  its threshold;
- texel sites lie in the site's tangent plane at radius-scaled offsets;
- the hand-written adjoint equals torch autograd of the forward (FP64) and
  FP64 central differences; a planted sign error in one derivative fails.
"""

import numpy as np
import pytest

from oracle import glue


def _params(seed, n=200, T=8):
    rng = np.random.default_rng(seed)
    return (rng.normal(0, 1, (n, 3)), rng.normal(0.01, 0.08, n), rng.normal(0.0, 0.1, n),
            rng.normal(0, 1, (n, 4)) * rng.uniform(0.5, 2.0, (n, 1)), rng.normal(0, 0.1, (n, T, 2)),
            rng.normal(0, 0.05, (n, T)))


def test_frame_is_scipy_rotation():
    from scipy.spatial.transform import Rotation
    p = _params(0)
    out = glue.forward(*p)
    q = p[3] / np.linalg.norm(p[3], axis=1, keepdims=True)
    R = Rotation.from_quat(np.concatenate([q[:, 1:], q[:, :1]], 1)).as_matrix()  # scipy takes (x, y, z, w)
    h, qn, u, ln = glue._frame(p[3])
    v = u / ln
    np.testing.assert_allclose(v[:, 0], R[:, 0, :], atol=1e-12)
    np.testing.assert_allclose(v[:, 1], R[:, 1, :], atol=1e-12)
    np.testing.assert_allclose(v[:, 2], R[:, 2, :], atol=1e-12)
    np.testing.assert_allclose(np.cross(v[:, 0], v[:, 1]), v[:, 2], atol=1e-12)
    np.testing.assert_allclose(out["normals"], R[:, 0, :], atol=1e-12)


def test_softplus_is_torchs():
    import torch
    x = np.concatenate([np.linspace(-0.5, 0.5, 1001), [0.19, 0.2, 0.2000001, 0.21, 3.0]])
    want = torch.nn.functional.softplus(torch.tensor(x, dtype=torch.float64), beta=100).numpy()
    np.testing.assert_allclose(glue.softplus(x), want, rtol=1e-12, atol=1e-15)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    torch.nn.functional.softplus(xt, beta=100).sum().backward()
    np.testing.assert_allclose(glue.softplus_grad(x), xt.grad.numpy(), rtol=1e-12, atol=1e-15)


def test_texel_sites_lie_in_the_tangent_plane():
    p = _params(1)
    out = glue.forward(*p)
    d = out["texel_sites"] - p[0][:, None, :]
    np.testing.assert_allclose((d * out["normals"][:, None, :]).sum(-1), 0.0, atol=1e-12)
    want = out["radii"][:, None] * np.linalg.norm(p[4], axis=-1)
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), want, rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(out["texel_height"], p[5] * out["radii"][:, None], rtol=1e-15)


def _torch_forward(t):
    import torch
    pts, rr, dr, q, off, hr = t
    r = torch.nn.functional.softplus(rr, beta=100)
    qh = q / q.norm(dim=-1, keepdim=True)
    w, x, y, z = qh.unbind(-1)
    nrm = torch.stack([1 - 2 * (y ** 2 + z ** 2), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1)
    tan = torch.stack([2 * (x * y + z * w), 1 - 2 * (x ** 2 + z ** 2), 2 * (y * z - x * w)], -1)
    bit = torch.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x ** 2 + y ** 2)], -1)
    nrm, tan, bit = (v / v.norm(dim=-1, keepdim=True) for v in (nrm, tan, bit))
    o = off * r[:, None, None]
    ts = pts[:, None, :] + o[..., 0:1] * tan[:, None, :] + o[..., 1:2] * bit[:, None, :]
    return dict(radii=r, density=torch.nn.functional.softplus(dr, beta=100), normals=nrm, texel_sites=ts,
                texel_height=hr * r[:, None])


def _grads_in(p, seed):
    rng = np.random.default_rng(seed)
    n, T = p[4].shape[:2]
    return dict(radii=rng.normal(size=n), density=rng.normal(size=n), normals=rng.normal(size=(n, 3)),
                texel_sites=rng.normal(size=(n, T, 3)), texel_height=rng.normal(size=(n, T)))


NAMES = ("points", "radii", "density", "quaternions", "texel_offsets", "texel_height")


def test_adjoint_matches_torch_autograd():
    import torch
    p = _params(2)
    g = _grads_in(p, 3)
    t = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in p]
    out = _torch_forward(t)
    ref = glue.forward(*p)
    for k in out:
        np.testing.assert_allclose(out[k].detach().numpy(), ref[k], rtol=1e-12, atol=1e-14)
    sum((out[k] * torch.tensor(g[k])).sum() for k in out).backward()
    got = glue.backward(*p, g)
    for name, a in zip(NAMES, t):
        np.testing.assert_allclose(got[name], a.grad.numpy(), rtol=1e-9, atol=1e-12, err_msg=name)


def test_adjoint_matches_central_differences():
    p = [np.array(a, np.float64) for a in _params(4, n=20, T=3)]
    g = _grads_in(p, 5)
    loss = lambda pp: sum((glue.forward(*pp)[k] * g[k]).sum() for k in g)
    got = glue.backward(*p, g)
    rng = np.random.default_rng(6)
    eps = 1e-6
    for i, name in enumerate(NAMES):
        flat = p[i].reshape(-1)
        for j in rng.choice(flat.size, min(flat.size, 12), replace=False):
            old = flat[j]
            flat[j] = old + eps
            lp = loss(p)
            flat[j] = old - eps
            lm = loss(p)
            flat[j] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(got[name].reshape(-1)[j] - fd) < 1e-6 * (1 + abs(fd)), (name, j, got[name].reshape(-1)[j], fd)


def test_planted_sign_error_is_caught(monkeypatch):
    import torch
    p = _params(7, n=30)
    g = _grads_in(p, 8)
    t = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in p]
    out = _torch_forward(t)
    sum((out[k] * torch.tensor(g[k])).sum() for k in out).backward()
    real = glue.softplus_grad
    monkeypatch.setattr(glue, "softplus_grad", lambda x: -real(x))
    bad = glue.backward(*p, g)
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(bad["radii"], t[1].grad.numpy(), rtol=1e-9, atol=1e-12)
