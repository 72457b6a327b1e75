"""ORACLE -- test infrastructure only: FP64 numpy restatement of the
reference's Spherical-Voronoi texel colour head (powerfoam/color_fn.py).

  forward  color_fn.py:54-101   d = normalize(p - eye); grey 0.5 (and weight
           sum 0) when d . forward < cutoff; else w_i = exp(-temp_i |d - a_i|),
           rgb = max(0, sum w_i val_i / sum w_i + 0.5)
  backward color_fn.py:105-184  dL/drgb zeroed where rgb <= 0; per lobe with
           w_i / sum w >= 1e-3 (exact=False; all lobes with exact=True):
           dL/dval_i = w_i g / sum w, with c_i = (g . val_i - g . (rgb - 0.5)) / sum w:
           dL/daxis_i = temp_i w_i (d - a_i) / |d - a_i| c_i,
           dL/dtemp_i = -|d - a_i| w_i c_i
  cutoff   color_fn.py:218-222  cos(atan(sqrt(|right|^2 + |up|^2)) + 10 deg)
  layout   scene.py:290-312     lobe-major [D, P, c]; axis = raw / |raw|,
           temp = |raw| from the [N, T, 3 D] parameters

Pinned by tests/test_oracle_sv.py: closed forms (one lobe, identical lobes,
the cut-off cone), the exact adjoint against torch autograd and FP64
central differences, and the 1e-3 skip against the exact adjoint.
"""

import math

import numpy as np


def fov_cos_cutoff(right, up):
    tan_diag = math.sqrt(float(np.dot(right, right)) + float(np.dot(up, up)))
    return math.cos(math.atan(tan_diag) + math.radians(10.0))


def camera_forward(right, up):
    f = np.cross(np.asarray(up, np.float64), np.asarray(right, np.float64))
    return f / np.linalg.norm(f)


def att_from_params(sv_axis, sv_rgb, dof):
    """scene.py:290-312: [N,T,3D] raw axes and values -> lobe-major axis
    [D,NT,3], values [D,NT,3], temps [D,NT]."""
    ax = np.asarray(sv_axis, np.float64).reshape(-1, sv_axis.shape[1], dof, 3)
    temp = np.linalg.norm(ax, axis=-1)
    ax = ax / temp[..., None]
    val = np.asarray(sv_rgb, np.float64).reshape(-1, sv_rgb.shape[1], dof, 3)
    return (ax.transpose(2, 0, 1, 3).reshape(dof, -1, 3), val.transpose(2, 0, 1, 3).reshape(dof, -1, 3),
            temp.transpose(2, 0, 1).reshape(dof, -1))


def _dirs(points, eye):
    v = np.asarray(points, np.float64) - np.asarray(eye, np.float64)[None]
    ln = np.linalg.norm(v, axis=-1, keepdims=True)
    return np.where(ln > 0, v / np.where(ln > 0, ln, 1), 0.0)


def forward(points, eye, fwd, cutoff, axis, val, temp):
    """(rgb [P,3], weights_sum [P])."""
    d = _dirs(points, eye)
    cut = d @ np.asarray(fwd, np.float64) < cutoff
    e = d[None] - np.asarray(axis, np.float64)
    dist = np.linalg.norm(e, axis=-1)
    w = np.exp(-np.asarray(temp, np.float64) * dist)
    ws = w.sum(0)
    vs = (w[..., None] * np.asarray(val, np.float64)).sum(0)
    rgb = np.maximum(vs / ws[:, None] + 0.5, 0.0)
    rgb[cut] = 0.5
    ws = np.where(cut, 0.0, ws)
    return rgb, ws


def backward(points, eye, fwd, cutoff, axis, val, temp, rgb, ws, grad, exact=False):
    """(grad axis [D,P,3], grad val [D,P,3], grad temp [D,P])."""
    d = _dirs(points, eye)
    cut = d @ np.asarray(fwd, np.float64) < cutoff
    g = np.where(rgb <= 0.0, 0.0, np.asarray(grad, np.float64))
    dp_out = (g * (rgb - 0.5)).sum(-1)
    e = d[None] - np.asarray(axis, np.float64)
    dist = np.linalg.norm(e, axis=-1)
    w = np.exp(-np.asarray(temp, np.float64) * dist)
    wsafe = np.where(cut, 1.0, ws)
    keep = ~cut[None] & (True if exact else ~(w / wsafe[None] < 1e-3))
    c = ((g[None] * np.asarray(val, np.float64)).sum(-1) - dp_out[None]) / wsafe[None]
    ga = np.where(keep[..., None], e * (np.asarray(temp) * w / dist * c)[..., None], 0.0)
    gv = np.where(keep[..., None], g[None] * (w / wsafe[None])[..., None], 0.0)
    gt = np.where(keep, -dist * w * c, 0.0)
    return ga, gv, gt
