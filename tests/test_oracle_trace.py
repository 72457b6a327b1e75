"""Pins of the oracle's ray tracer (raytrace.py:115-241 plus depth, normals,
cell lists and an exact adjoint):

1. the walked cell sequence equals the power cells the ray actually crosses,
   found independently in FP64 (the ray is cut at every radical plane and the
   power-argmin site owns each piece);
2. with exact geometry and no early stop, colour, opacity and normals equal
   the exact volume renderer of the scene definition;
3. the adjoint matches FP64 central finite differences.
"""

import numpy as np
import pytest

from oracle import orc
from powerfoam_amd import synth
from tests.test_oracle_raster import PARAMS, camera, exact_render, inputs, make_scene


def trace_scene(seed, n=30, texel_mode="uniform", height_std=0.0):
    s = make_scene(seed, n=n, texel_mode=texel_mode, height_std=height_std)
    # full power (weighted Delaunay) adjacency, not alpha-filtered
    s["adj"], s["off"] = synth.power_adjacency(s["points"], s["radii"], alpha_complex=False)
    return s


def start_cell(s, eye):
    pw = ((s["points"].astype(np.float64) - np.asarray(eye)) ** 2).sum(1) - s["radii"].astype(np.float64) ** 2
    return int(np.argmin(pw))


def exact_cells(s, eye, d):
    """Power cells crossed by the ray eye + t d, t >= 0, in order (FP64)."""
    p = s["points"].astype(np.float64)
    r = s["radii"].astype(np.float64)
    w = (p * p).sum(1) - r * r
    dc = p[None] - p[:, None]
    rhs = w[None] - w[:, None]
    den = 2 * dc @ d
    num = rhs - 2 * dc @ eye
    with np.errstate(divide="ignore", invalid="ignore"):
        tp = num / den
    ts = np.unique(np.concatenate([[0.0], tp[np.isfinite(tp) & (tp > 0)].ravel(), [1e3]]))
    seq = []
    for a, b in zip(ts[:-1], ts[1:]):
        if b - a < 1e-9:
            continue
        x = eye + 0.5 * (a + b) * d
        k = int(np.argmin(((x - p) ** 2).sum(1) - r * r))
        if not seq or seq[-1] != k:
            seq.append(k)
    return seq


@pytest.mark.parametrize("seed,eye", [(0, (0.2, 0.1, 2.4)), (1, (-1.5, 0.8, 1.6))])
def test_walk_visits_the_power_cells_along_the_ray(seed, eye):
    s = trace_scene(seed)
    t, cam = camera(24, 18, eye)
    sp, inp = inputs(s, exact_geometry=True)
    out = orc.trace_fwd(cam, inp, start_cell(s, eye), threshold=0.0, record_cells=64)
    right, up = t.right.numpy().astype(np.float64), t.up.numpy().astype(np.float64)
    fwd = np.cross(up, right)
    fwd /= np.linalg.norm(fwd)
    e = np.asarray(eye, np.float64)
    checked = 0
    for i in range(t.height):
        for j in range(t.width):
            x = 2.0 * j / (t.width - 1) - 1.0
            y = 1.0 - 2.0 * i / (t.height - 1)
            d = x * right + y * up + fwd
            d /= np.linalg.norm(d)
            want = exact_cells(s, e, d)
            n = out["n_cells"][i, j]
            got = out["cells"][i, j, :n].tolist()
            assert got == want, (i, j, got, want)
            checked += len(want) > 1
    assert checked > 100


@pytest.mark.parametrize("seed,eye", [(0, (0.2, 0.1, 2.4)), (2, (0.0, -2.0, 1.2))])
def test_trace_matches_exact_volume_renderer(seed, eye):
    s = trace_scene(seed)
    t, cam = camera(36, 28, eye)
    sp, inp = inputs(s, exact_geometry=True)
    out = orc.trace_fwd(cam, inp, start_cell(s, eye), threshold=0.0)
    color, opac, normal = exact_render(s, t)
    assert opac.max() > 0.5
    np.testing.assert_allclose(1 - np.exp(out["log_t"]), opac, atol=1e-8)
    np.testing.assert_allclose(out["color"], color, atol=1e-8)
    np.testing.assert_allclose(out["normal"], normal, atol=1e-8)


def _trace_objective(s, cam, eye, w):
    sp, inp = inputs(s, exact_geometry=True)
    out = orc.trace_fwd(cam, inp, start_cell(s, eye), threshold=0.0)
    valid = out["depth"] != -1
    L = (out["color"] * w["color"]).sum() + (out["normal"] * w["normal"]).sum()
    L += ((1 - np.exp(out["log_t"])) * w["opacity"]).sum() + (np.where(valid, out["depth"], 0) * w["depth"]).sum()
    return L, out, inp


@pytest.mark.parametrize("mode", ["uniform", "varied"])
def test_trace_adjoint_matches_central_differences(mode):
    s = trace_scene(21 if mode == "uniform" else 22, texel_mode=mode, height_std=0.0 if mode == "uniform" else 0.1)
    eye = (0.2, 0.1, 2.4)
    t, cam = camera(30, 22, eye)
    rng = np.random.default_rng(5)
    H, W = t.height, t.width
    w = dict(color=rng.normal(size=(H, W, 3)), normal=rng.normal(size=(H, W, 3)), opacity=rng.normal(size=(H, W)),
             depth=rng.normal(size=(H, W)))
    L0, out, inp = _trace_objective(s, cam, eye, w)
    wd = np.where(out["depth"] != -1, w["depth"], 0.0)
    g = orc.trace_bwd(cam, inp, start_cell(s, eye), out, w["color"], w["normal"], wd, w["opacity"], threshold=0.0)
    keys = [k for k in PARAMS if not (mode == "varied" and k == "radii")]
    if mode == "uniform":
        keys = [k for k in keys if k not in ("texel_rgb", "texel_height", "texel_sites")]
    tested = 0
    for trial in range(256):
        k = keys[trial % len(keys)]
        arr = s[k]
        idx = tuple(rng.integers(0, d) for d in arr.shape)
        gk, col = PARAMS[k]
        site = idx[0]
        if k in ("points", "normals"):
            ana = g[gk][site, idx[1]]
        elif k in ("radii", "density"):
            ana = g[gk][site, col]
        else:
            ana = g[gk][idx]
        base = float(arr[idx])
        fds = []
        for eps in (1e-4, 1e-5):
            arr[idx] = base + eps
            xp = float(arr[idx])
            Lp, _, _ = _trace_objective(s, cam, eye, w)
            arr[idx] = base - eps
            xm = float(arr[idx])
            Lm, _, _ = _trace_objective(s, cam, eye, w)
            arr[idx] = base
            fds.append((Lp - Lm) / (xp - xm))
        if abs(fds[0] - fds[1]) > 1e-3 * (abs(fds[0]) + 1.0):
            continue
        tested += 1
        assert abs(ana - fds[0]) <= 2e-3 * abs(fds[0]) + 2e-3, (k, idx, ana, fds)
    assert tested >= 200, tested


def test_fragile_rays_carry_the_precision_sensitivity():
    """The rays the FP64 walk marks fragile are the ones whose gradients move
    under float32 rounding: with them weighted out, the FP32 oracle (which
    keeps the reference's |oc|^2 - r^2 discriminant) agrees with FP64 within
    1e-3; with them in, it does not."""
    from powerfoam_amd.camera import look_at
    from tests import harness
    s = synth.scene_uniform(3, n=1500, n_texels=4, r_range=(0.03, 0.12), height_std=0.05)
    s["texel_rgb"] = np.random.default_rng(3).uniform(0, 1, (1500, 4, 3)).astype(np.float32)
    s["adj"], s["off"] = synth.power_adjacency(s["points"], s["radii"], alpha_complex=False)
    t = look_at((0.1, 0.2, 2.2), (0.0, 0.0, 0.0), 40, 32, 60.0)
    cam, _k = orc.camera(t.eye.numpy(), t.right.numpy(), t.up.numpy(), t.width, t.height)
    sp, inp = inputs(s, exact_geometry=True)
    st = start_cell(s, t.eye.numpy().astype(np.float64))
    rng = np.random.default_rng(7)
    H, W = t.height, t.width
    w = dict(color=rng.normal(size=(H, W, 3)), normal=rng.normal(size=(H, W, 3)), opacity=rng.normal(size=(H, W)),
             depth=rng.normal(size=(H, W)))
    f64 = orc.trace_fwd(cam, inp, st)
    assert 0.01 < f64["fragile"].mean() < 0.5
    errs = {}
    for masked in (False, True):
        m = ~f64["fragile"] if masked else np.ones((H, W), bool)
        g = {}
        for p in ("f64", "f32"):
            ref = orc.trace_fwd(cam, inp, st, precision=p)
            wd = np.where(ref["depth"] != -1, w["depth"], 0.0) * m
            g[p] = orc.trace_bwd(cam, inp, st, ref, w["color"] * m[..., None], w["normal"] * m[..., None], wd,
                                 w["opacity"] * m, precision=p)
        errs[masked] = max(harness.rel_err(g["f32"][k], g["f64"][k]) for k in harness.GRAD_KEYS)
    assert errs[True] < 1e-3, errs
    assert errs[False] > 3 * errs[True], errs
