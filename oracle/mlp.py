"""ORACLE -- test infrastructure only: the CPU restatement of the learned
colour head (powerfoam_amd/csrc/mlp.hip) that tests/ use as the checker; the
product never imports it. The head has no reference counterpart: it fills
the texel-colour slot of the reference's SphericalVoronoi pre-pass
(color_fn.py:54-184), with the encoding e(d) = (1, Y_1(d) .. Y_15(d)) of the
SH basis (sh_utils.py:4-49), or the 16-feature frequency encoding
(1, d, sin(pi d), cos(pi d), sin(2 pi d), cos(2 pi d)).

numpy, FP64 arithmetic. With rounding=True it rounds to bf16 (nearest-even)
at the points the kernels do (inputs, weights, h1, h2; in the backward the
output gradient and dZ2, dZ1 before their products and before the bias
gradients' sums, as a bf16 linear layer's backward sums its bf16 output
gradient); with rounding=False it is the plain FP64 network whose
derivatives the finite-difference tests pin.
"""

import numpy as np

from powerfoam_amd.synth import sh_basis


def bf16(x):
    """Round to bfloat16, nearest-even (finite inputs), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(d, encoding="sh"):
    """e(d) at unit directions d [...,3]: (1, Y_1..Y_15) for "sh"; for "freq"
    (1, d, sin(pi d), cos(pi d), sin(2 pi d), cos(2 pi d)), per component."""
    d = np.asarray(d, np.float64)
    one = np.ones(d.shape[:-1] + (1,))
    if encoding == "freq":
        a = np.pi * d
        return np.concatenate([one, d, np.sin(a), np.cos(a), np.sin(2 * a), np.cos(2 * a)], -1)
    return np.concatenate([one, sh_basis(d)], -1)


def rows_texels(texel_sites, eye, T):
    """(site index, unit direction) per texel row (color_fn.py:72)."""
    ts = np.asarray(texel_sites, np.float32).reshape(-1, 3)
    d = ts.astype(np.float32) - np.asarray(eye, np.float32)[None]
    d = d.astype(np.float64)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return np.arange(ts.shape[0]) // T, d


def forward(params, sites, dirs, rounding=True, encoding="sh"):
    """rgb [R,3] and the saved activations. params: dict of latents [N,16],
    w1 [64,32], b1, w2 [64,64], b2, w3 [3,64], b3 (any float dtype)."""
    q = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    z = q(params["latents"])[sites]
    e = q(encode(dirs, encoding))
    x = np.concatenate([z, e], -1)
    w1, w2, w3 = q(params["w1"]), q(params["w2"]), q(params["w3"])
    b1, b2, b3 = (np.asarray(params[k], np.float64) for k in ("b1", "b2", "b3"))
    z1 = x @ w1.T + b1
    h1 = q(np.maximum(z1, 0.0))
    z2 = h1 @ w2.T + b2
    h2 = q(np.maximum(z2, 0.0))
    o = h2 @ w3.T + b3
    saved = dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, o=o, w1=w1, w2=w2, w3=w3)
    return np.maximum(o, 0.0), saved


def backward(params, sites, dirs, grad_rgb, rounding=True, encoding="sh"):
    """Gradients of sum(grad_rgb * rgb) w.r.t. latents and weights (dict)."""
    q = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    _, s = forward(params, sites, dirs, rounding, encoding)
    g = np.asarray(grad_rgb, np.float64)
    d_o = np.where(s["o"] > 0.0, g, 0.0)
    d_ob = q(d_o)
    dz2 = (d_ob @ s["w3"]) * (s["z2"] > 0.0)
    dz2b = q(dz2)
    dz1 = (dz2b @ s["w2"]) * (s["z1"] > 0.0)
    dz1b = q(dz1)
    dx = dz1b @ s["w1"]
    n = np.asarray(params["latents"]).shape[0]
    g_lat = np.zeros((n, 16))
    np.add.at(g_lat, np.asarray(sites), dx[:, :16])
    # bias gradients sum the same (rounded) dZ / dOut the weight gradients use
    return dict(latents=g_lat, w1=dz1b.T @ s["x"], b1=dz1b.sum(0), w2=dz2b.T @ s["h1"], b2=dz2b.sum(0),
                w3=d_ob.T @ s["h2"], b3=d_ob.sum(0))
