"""ORACLE -- test infrastructure only: FP64 numpy restatement of the
reference's scene glue (powerfoam/scene.py:239-337), the map from trainable
parameters to the rasterizer's per-site inputs, with a hand-written adjoint.

  radii   = softplus(radii_raw, beta=100)          scene.py:287-288
  density = softplus(density_raw, beta=100)        scene.py:239-240
  q^ = q / |q|                                     scene.py:243, 260
  normal    = normalize(1-2(y^2+z^2), 2(xy-zw), 2(xz+yw))     scene.py:248-256
  tangent   = normalize(2(xy+zw), 1-2(x^2+z^2), 2(yz-xw))     scene.py:265-273
  bitangent = normalize(2(xz-yw), 2(yz+xw), 1-2(x^2+y^2))     scene.py:276-284
  texel_sites = points + (off0 r) tangent + (off1 r) bitangent  scene.py:321-329
  texel_height = height_raw r                      scene.py:337

softplus as torch computes it: x where beta x > 20, else log1p(exp(beta x))/beta.
Pinned by tests/test_oracle_glue.py: the frame against scipy's rotation
matrix of the same quaternion, softplus against torch's, the adjoint against
torch autograd and FP64 central differences.
"""

import numpy as np

BETA = 100.0
THRESHOLD = 20.0


def softplus(x):
    bx = BETA * x
    return np.where(bx > THRESHOLD, x, np.log1p(np.exp(np.minimum(bx, THRESHOLD))) / BETA)


def softplus_grad(x):
    bx = BETA * x
    return np.where(bx > THRESHOLD, 1.0, 1.0 / (1.0 + np.exp(-np.minimum(bx, THRESHOLD))))


def _frame(q):
    qn = np.linalg.norm(q, axis=-1, keepdims=True)
    h = q / qn
    w, x, y, z = h[:, 0], h[:, 1], h[:, 2], h[:, 3]
    u = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)  # [N, 3 (normal, tangent, bitangent), 3]
    ln = np.linalg.norm(u, axis=-1, keepdims=True)
    return h, qn, u, ln


def forward(points, radii_raw, density_raw, quaternions, texel_offsets, height_raw):
    p = np.asarray(points, np.float64)
    q = np.asarray(quaternions, np.float64)
    off = np.asarray(texel_offsets, np.float64)
    r = softplus(np.asarray(radii_raw, np.float64))
    h, qn, u, ln = _frame(q)
    v = u / ln
    ts = p[:, None, :] + (off[..., 0:1] * r[:, None, None]) * v[:, None, 1, :] + \
        (off[..., 1:2] * r[:, None, None]) * v[:, None, 2, :]
    return dict(radii=r, density=softplus(np.asarray(density_raw, np.float64)), normals=v[:, 0],
                texel_sites=ts, texel_height=np.asarray(height_raw, np.float64) * r[:, None])


def backward(points, radii_raw, density_raw, quaternions, texel_offsets, height_raw, g):
    """Gradients of the six parameter arrays for incoming output gradients
    g (dict with any of radii, density, normals, texel_sites, texel_height);
    `points` gets the texel sites' share only."""
    q = np.asarray(quaternions, np.float64)
    off = np.asarray(texel_offsets, np.float64)
    hr = np.asarray(height_raw, np.float64)
    rr = np.asarray(radii_raw, np.float64)
    n, T = off.shape[:2]
    r = softplus(rr)
    h, qn, u, ln = _frame(q)
    v = u / ln
    z3 = lambda *s: np.zeros(s)
    gr = np.asarray(g.get("radii", z3(n)), np.float64)
    gd = np.asarray(g.get("density", z3(n)), np.float64)
    gn = np.asarray(g.get("normals", z3(n, 3)), np.float64)
    gx = np.asarray(g.get("texel_sites", z3(n, T, 3)), np.float64)
    gh = np.asarray(g.get("texel_height", z3(n, T)), np.float64)
    gxt = (gx * v[:, None, 1]).sum(-1)
    gxb = (gx * v[:, None, 2]).sum(-1)
    g_r = gr + (off[..., 0] * gxt + off[..., 1] * gxb + gh * hr).sum(1)
    gv = np.stack([gn, (gx * (off[..., 0:1] * r[:, None, None])).sum(1),
                   (gx * (off[..., 1:2] * r[:, None, None])).sum(1)], 1)
    gu = (gv - v * (v * gv).sum(-1, keepdims=True)) / ln
    w, x, y, z = h[:, 0], h[:, 1], h[:, 2], h[:, 3]
    o = np.zeros_like(w)
    # d(row c of R)/d(w, x, y, z): [N, 3 rows, 4 params, 3]
    d = np.stack([
        np.stack([np.stack([o, -2 * z, 2 * y], -1), np.stack([o, 2 * y, 2 * z], -1),
                  np.stack([-4 * y, 2 * x, 2 * w], -1), np.stack([-4 * z, -2 * w, 2 * x], -1)], 1),
        np.stack([np.stack([2 * z, o, -2 * x], -1), np.stack([2 * y, -4 * x, -2 * w], -1),
                  np.stack([2 * x, o, 2 * z], -1), np.stack([2 * w, -4 * z, 2 * y], -1)], 1),
        np.stack([np.stack([-2 * y, 2 * x, o], -1), np.stack([2 * z, 2 * w, -4 * x], -1),
                  np.stack([-2 * w, 2 * z, -4 * y], -1), np.stack([2 * x, 2 * y, o], -1)], 1),
    ], 1)
    gqh = np.einsum("ncj,ncpj->np", gu, d)
    gq = (gqh - h * (h * gqh).sum(-1, keepdims=True)) / qn
    return dict(points=gx.sum(1), radii=g_r * softplus_grad(rr), density=gd * softplus_grad(np.asarray(
        density_raw, np.float64)), quaternions=gq, texel_offsets=np.stack([r[:, None] * gxt, r[:, None] * gxb], -1),
        texel_height=gh * r[:, None])
