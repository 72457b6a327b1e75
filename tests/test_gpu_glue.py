"""GPU scene glue (pf_glue_fwd / pf_glue_bwd, scene.py:239-337) against the
FP64 oracle (oracle/glue.py): outputs within 2e-6 relative, each parameter
gradient within 1e-5 relative (norm); and the glue chained into the
rasterizer: gradients of the raw parameters through both equal the
rasterizer's gradients pushed through the oracle's adjoint."""

import numpy as np
import pytest

from oracle import glue as og
from tests import harness

pytestmark = pytest.mark.gpu


def _params(seed, n=3001, T=8):
    rng = np.random.default_rng(seed)
    return (rng.normal(0, 1, (n, 3)), rng.normal(0.0, 0.08, n), rng.normal(0.0, 0.1, n),
            rng.normal(0, 1, (n, 4)), rng.normal(0, 0.1, (n, T, 2)), rng.normal(0, 0.05, (n, T)))


def test_glue_forward_backward_match_oracle():
    import torch
    from powerfoam_amd.scene import GlueFn
    p = [a.astype(np.float32) for a in _params(0)]
    print("glue fwd/bwd n=3001 T=8", flush=True)
    t = [torch.tensor(a, device="cuda", requires_grad=True) for a in p]
    out = GlueFn.apply(*t)
    ref = og.forward(*p)
    for k, v in zip(("radii", "density", "normals", "texel_sites", "texel_height"), out):
        err = harness.rel_err(v.detach().cpu().numpy(), ref[k])
        print(k, err, flush=True)
        assert err < 2e-6, (k, err)
    rng = np.random.default_rng(1)
    g = {k: rng.normal(size=ref[k].shape).astype(np.float32) for k in ref}
    sum((v * torch.tensor(g[k], device="cuda")).sum()
        for k, v in zip(("radii", "density", "normals", "texel_sites", "texel_height"), out)).backward()
    want = og.backward(*p, g)
    for name, a in zip(("points", "radii", "density", "quaternions", "texel_offsets", "texel_height"), t):
        err = harness.rel_err(a.grad.cpu().numpy(), want[name])
        print(name, err, flush=True)
        assert err < 1e-5, (name, err)


def test_glue_with_missing_output_gradients():
    """Outputs the loss does not use come back as None; their gradient terms are zero."""
    import torch
    from powerfoam_amd.scene import GlueFn
    p = [a.astype(np.float32) for a in _params(2, n=500)]
    t = [torch.tensor(a, device="cuda", requires_grad=True) for a in p]
    out = GlueFn.apply(*t)
    g = np.random.default_rng(3).normal(size=p[4].shape[:2] + (3,)).astype(np.float32)
    (out[3] * torch.tensor(g, device="cuda")).sum().backward()
    want = og.backward(*p, dict(texel_sites=g))
    for name, a in zip(("points", "radii", "density", "quaternions", "texel_offsets", "texel_height"), t):
        assert harness.rel_err(a.grad.cpu().numpy(), want[name]) < 1e-5 or np.abs(want[name]).max() == 0, name
    assert np.all(t[2].grad.cpu().numpy() == 0)


def test_glue_into_rasterizer_chain():
    import torch
    from powerfoam_amd import synth
    from powerfoam_amd.camera import look_at
    from powerfoam_amd.rasterize import Rasterizer
    from powerfoam_amd.scene import SceneGlue
    n, T = 600, 8
    rng = np.random.default_rng(5)
    pts = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    r_act = np.exp(rng.uniform(np.log(0.03), np.log(0.15), n))
    radii_raw = (r_act + np.log(-np.expm1(-100 * r_act)) / 100).astype(np.float32)  # softplus^-1
    dens_raw = rng.uniform(0.02, 0.3, n).astype(np.float32)
    q = rng.normal(size=(n, 4)).astype(np.float32)
    off = rng.normal(0, 0.1, (n, T, 2)).astype(np.float32)
    hr = rng.normal(0, 0.05, (n, T)).astype(np.float32)
    sg = SceneGlue(pts, radii_raw, q, dens_raw, off, hr, device="cuda")
    x = sg.rasterizer_inputs()
    adj, off_csr = synth.cech_adjacency(pts, x["radii"].detach().cpu().numpy())
    cam = look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 64, 48, 60.0)
    rgb = torch.tensor(rng.uniform(0, 1, (n, T, 3)).astype(np.float32), device="cuda")
    r = Rasterizer(harness._args(T), "cuda")
    print("glue into rasterizer 600 sites 64x48", flush=True)
    out = r.forward(cam, None, x["points"], x["radii"], x["density"], x["normals"], x["texel_sites"], rgb,
                    x["texel_height"], torch.tensor(adj, device="cuda"), torch.tensor(off_csr, device="cuda"), None,
                    False)
    gc = torch.tensor(rng.normal(size=(48, 64, 3)).astype(np.float32), device="cuda")
    # the rasterizer's gradients w.r.t. its inputs, then through the oracle glue
    leaves = {k: v.detach().requires_grad_(True) for k, v in x.items()}
    out2 = r.forward(cam, None, leaves["points"], leaves["radii"], leaves["density"], leaves["normals"],
                     leaves["texel_sites"], rgb, leaves["texel_height"], torch.tensor(adj, device="cuda"),
                     torch.tensor(off_csr, device="cuda"), None, False)
    (out2[0] * gc).sum().backward()
    (out[0] * gc).sum().backward()
    gin = {k: leaves[k].grad.cpu().numpy() for k in ("radii", "density", "normals", "texel_sites", "texel_height")}
    pin = [a.detach().cpu().numpy() for a in (sg.points, sg.radii, sg.density, sg.quaternions, sg.texel_sites,
                                              sg.texel_height)]
    want = og.backward(*pin, gin)
    want["points"] = want["points"] + leaves["points"].grad.cpu().numpy()
    got = dict(points=sg.points.grad, radii=sg.radii.grad, density=sg.density.grad, quaternions=sg.quaternions.grad,
               texel_offsets=sg.texel_sites.grad, texel_height=sg.texel_height.grad)
    for k, v in got.items():
        err = harness.rel_err(v.cpu().numpy(), want[k])
        print(k, err, flush=True)
        assert err < 1e-4, (k, err)
    assert np.abs(want["radii"]).max() > 0 and np.abs(want["quaternions"]).max() > 0
