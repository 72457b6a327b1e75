"""GPU Spherical-Voronoi colour head (pf_sv_fwd / pf_sv_bwd, color_fn.py:54-184)
against the FP64 oracle (oracle/sv.py), and a trained-scene file in the
reference's model.pt layout rendered through the drop-in path (reader, scene
glue, SV colour, rasterizer) against the oracle chain (oracle glue, oracle
SV, FP64 rasterizer), forward and backward."""

import os
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import glue as og
from oracle import orc
from oracle import sv as osv
from tests import harness

pytestmark = pytest.mark.gpu


def _cam():
    from powerfoam_amd.camera import look_at
    return look_at((0.3, 0.2, 3.0), (0.0, 0.0, 0.0), 64, 48, 60.0)


def test_sv_forward_backward_match_oracle():
    import torch
    from powerfoam_amd.color_fn import SphericalVoronoi
    rng = np.random.default_rng(0)
    P, D = 5001, 8
    pts = rng.uniform(-1.5, 1.5, (P, 3)).astype(np.float32)
    ax = rng.normal(size=(D, P, 3))
    temp = (np.linalg.norm(ax, axis=-1) * 2.0).astype(np.float32)
    ax = (ax / np.linalg.norm(ax, axis=-1, keepdims=True)).astype(np.float32)
    val = rng.normal(0, 0.4, (D, P, 3)).astype(np.float32)
    cam = _cam()
    head = SphericalVoronoi(SimpleNamespace(sv_dof=D), "cuda")
    t = [torch.tensor(a, device="cuda", requires_grad=True) for a in (ax, val, temp)]
    print("sv fwd/bwd P=5001 D=8", flush=True)
    out = head.forward(torch.tensor(pts, device="cuda"), cam, *t)
    right, up, eye = (x.numpy().astype(np.float64) for x in (cam.right, cam.up, cam.eye))
    fwd, cut = osv.camera_forward(right, up), osv.fov_cos_cutoff(right, up)
    rgb, ws = osv.forward(pts, eye, fwd, cut, ax, val, temp)
    err = np.abs(out.detach().cpu().numpy() - rgb).max()
    print("fwd max abs", err, "cut off", (ws == 0).mean(), flush=True)
    assert err < 2e-6
    assert 0.005 < (ws == 0).mean() < 0.9  # some texels lie outside the camera's cone
    g = rng.normal(size=(P, 3)).astype(np.float32)
    (out * torch.tensor(g, device="cuda")).sum().backward()
    want = osv.backward(pts, eye, fwd, cut, ax, val, temp, rgb, ws, g)
    for name, a, w in zip(("axis", "values", "temps"), t, want):
        e = harness.rel_err(a.grad.cpu().numpy(), w)
        print(name, e, flush=True)
        assert e < 1e-5, (name, e)


def test_model_pt_renders_like_the_oracle_chain(tmp_path):
    import torch
    from powerfoam_amd import synth
    from powerfoam_amd.checkpoint import PowerfoamCheckpoint, save_pt
    from powerfoam_amd.color_fn import SphericalVoronoi
    from powerfoam_amd.rasterize import Rasterizer
    n, T, D = 700, 8, 8
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    r_act = np.exp(rng.uniform(np.log(0.03), np.log(0.14), n))
    s = dict(points=pts, density=rng.uniform(0.01, 0.3, n), radii=r_act + np.log(-np.expm1(-100 * r_act)) / 100,
             quaternions=rng.normal(size=(n, 4)), texel_sites=rng.normal(0, 0.1, (n, T, 2)),
             texel_sv_axis=rng.normal(size=(n, T, 3 * D)) * 2.0, texel_sv_rgb=rng.normal(0, 0.3, (n, T, 3 * D)),
             texel_height=rng.normal(0, 0.05, (n, T)))
    s = {k: np.asarray(v, np.float32) for k, v in s.items()}
    g0 = og.forward(s["points"], s["radii"], s["density"], s["quaternions"], s["texel_sites"], s["texel_height"])
    adj, off = synth.cech_adjacency(s["points"], g0["radii"].astype(np.float32))
    path = os.path.join(tmp_path, "model.pt")
    save_pt(path, adjacency=adj, adjacency_offsets=off, **s)
    print("model.pt round trip 700 sites, render 64x48", flush=True)
    ck = PowerfoamCheckpoint.load(path, device="cuda", requires_grad=True)
    cam = _cam()
    x = ck.render_inputs(cam, SphericalVoronoi(SimpleNamespace(sv_dof=D), "cuda"))
    r = Rasterizer(harness._args(T), "cuda")
    out = r.forward(cam, None, x["points"], x["radii"], x["density"], x["normals"], x["texel_sites"],
                    x["texel_rgb"], x["texel_height"], x["adjacency"], x["adjacency_offsets"], None, False)
    gc = rng.normal(size=(48, 64, 3)).astype(np.float32)
    (out[0] * torch.tensor(gc, device="cuda")).sum().backward()

    # the oracle chain on the same stored parameters
    g64 = og.forward(s["points"], s["radii"], s["density"], s["quaternions"], s["texel_sites"], s["texel_height"])
    right, up, eye = (v.numpy().astype(np.float64) for v in (cam.right, cam.up, cam.eye))
    ax, val, temp = osv.att_from_params(s["texel_sv_axis"].astype(np.float64), s["texel_sv_rgb"].astype(np.float64),
                                        D)
    fwd, cut = osv.camera_forward(right, up), osv.fov_cos_cutoff(right, up)
    tsites = g64["texel_sites"].reshape(-1, 3)
    rgb, ws = osv.forward(tsites, eye, fwd, cut, ax, val, temp)
    case = harness.Case(s["points"], g64["radii"].astype(np.float32), g64["density"].astype(np.float32),
                        g64["normals"].astype(np.float32), g64["texel_sites"].astype(np.float32),
                        rgb.reshape(n, T, 3).astype(np.float32), g64["texel_height"].astype(np.float32), adj, off,
                        cam, np.zeros((48, 64, 3), np.float32))
    ocam, _ = case.pf_camera()
    offs, sites, _ = orc.bin_lists(ocam, case.spheres())
    inp = harness.oracle_inputs(case)
    ref = orc.raster_fwd(ocam, inp, offs, sites, precision="f64")
    err = np.abs(out[0].detach().cpu().numpy() - ref["color"]).max()
    print("colour max abs", err, flush=True)
    assert err < 1e-4
    z = np.zeros((48, 64), np.float32)
    rg = orc.raster_bwd(ocam, inp, offs, sites, ref["tile_early_stop"], ref["log_t"], gc, z, z,
                        np.zeros((48, 64, 3), np.float32), np.zeros(n, np.float32), precision="f64")
    keep = ~rg["fragile"]
    # SV lobe values: the texel colour gradient through the SV adjoint, back to [N,T,3D]
    _, gv, _ = osv.backward(tsites, eye, fwd, cut, ax, val, temp, rgb, ws, rg["grad_texel_rgb"].reshape(-1, 3))
    want = gv.reshape(D, n, T, 3).transpose(1, 2, 0, 3).reshape(n, T, 3 * D)
    got = ck.texel_sv_rgb.grad.cpu().numpy()
    e = harness.rel_err(got[keep], want[keep])
    print("texel_sv_rgb grad rel", e, "flagged", rg["fragile"].mean(), flush=True)
    assert e < 1e-3
    # raw density through the glue adjoint
    wd = og.backward(s["points"], s["radii"], s["density"], s["quaternions"], s["texel_sites"], s["texel_height"],
                     dict(density=rg["grad_nsigmas"][:, 3]))["density"]
    e = harness.rel_err(ck.glue.density.grad.cpu().numpy()[keep], wd[keep])
    print("raw density grad rel", e, flush=True)
    assert e < 1e-3
