"""GPU ray tracer (pf_trace_fwd / pf_trace_bwd) against the oracle:
visited cell sequences bit-exact against the FP32 oracle walk (pinhole and an
equidistant fisheye given as ray maps), colour / opacity / depth / normals
within 1e-4 of the FP64 oracle, gradients within 1e-3 of the FP64 adjoint and
against FP64 central differences."""

from types import SimpleNamespace

import numpy as np
import pytest

from oracle import orc
from powerfoam_amd import synth
from powerfoam_amd.camera import TorchCamera, fisheye_ray_maps, look_at
from tests import harness
from tests.test_oracle_raster import PARAMS, make_scene
from tests.test_oracle_trace import _trace_objective, start_cell, trace_scene

pytestmark = pytest.mark.gpu

ABS_TOL = 1e-4
GRAD_REL_TOL = 1e-3


def big_scene(seed, n=3000, T=8):
    s = synth.scene_uniform(seed, n=n, n_texels=T, r_range=(0.03, 0.12), height_std=0.05)
    rng = np.random.default_rng(seed)
    s["texel_rgb"] = rng.uniform(0, 1, (n, T, 3)).astype(np.float32)
    s["adj"], s["off"] = synth.power_adjacency(s["points"], s["radii"], alpha_complex=False)
    return s


def cameras(kind):
    eye = (0.1, 0.2, 2.2)
    if kind == "pinhole":
        return look_at(eye, (0.0, 0.0, 0.0), 72, 56, 60.0)
    # fisheye from inside the scene: rays leave the start cell in every direction
    eye = (0.1, 0.2, 0.4)
    rm = fisheye_ray_maps(eye, (0.0, 0.0, -1.0), 64)
    import torch
    t = look_at(eye, (0.0, 0.0, -1.0), 64, 64, 60.0)
    return TorchCamera(t.eye, t.right, t.up, 64, 64, torch.tensor(rm))


def _tensors(s, dev):
    import torch
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    return dict(points=tt(s["points"]), radii=tt(s["radii"]), density=tt(s["density"]), normals=tt(s["normals"]),
                texel_sites=tt(s["texel_sites"]), texel_rgb=tt(s["texel_rgb"]),
                texel_height=tt(s["texel_height"]), adj=tt(s["adj"], torch.int32), off=tt(s["off"], torch.int32))


def _tracer(T):
    from powerfoam_amd.raytrace import RayTracer
    return RayTracer(SimpleNamespace(num_texel_sites=T, is_pinhole=True), "cuda")


def _gpu_cam(cam):
    return TorchCamera(cam.eye, cam.right, cam.up, cam.width, cam.height,
                       None if cam.ray_maps is None else cam.ray_maps.cuda().contiguous())


def _oracle_cam(cam):
    rm = None if cam.ray_maps is None else cam.ray_maps.numpy()
    return orc.camera(cam.eye.numpy(), cam.right.numpy(), cam.up.numpy(), cam.width, cam.height, rm)


@pytest.mark.parametrize("kind", ["pinhole", "fisheye"])
@pytest.mark.parametrize("encoded", [True, False], ids=["fp16faces", "exact"])
def test_cell_sequences_bitexact(kind, encoded):
    import torch
    from powerfoam_amd.rasterize import power_face_encoding
    s = big_scene(1)
    cam = cameras(kind)
    x = _tensors(s, "cuda")
    diff = power_face_encoding(x["points"], x["radii"], x["adj"], x["off"]) if encoded else None
    st = start_cell(s, cam.eye.numpy().astype(np.float64))
    print(f"cells {kind} encoded={encoded} start={st}", flush=True)
    cells, n = _tracer(8).trace_cells(_gpu_cam(cam), st, x["points"], x["radii"], x["density"], x["normals"],
                                      x["texel_sites"], x["texel_rgb"], x["texel_height"], x["adj"], x["off"], diff,
                                      0.0, max_cells=128)
    ocam, _k = _oracle_cam(cam)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    ns = np.concatenate([s["normals"], s["density"][:, None]], 1)
    rgbh = np.concatenate([s["texel_rgb"], s["texel_height"][..., None]], -1)
    inp = orc.Inputs(sp, ns, s["texel_sites"], rgbh, s["adj"], s["off"],
                     None if diff is None else diff.cpu().numpy().view(np.uint16))
    ref = orc.trace_fwd(ocam, inp, st, threshold=0.0, record_cells=128, precision="f32")
    assert np.array_equal(n.cpu().numpy(), ref["n_cells"])
    assert np.array_equal(cells.cpu().numpy(), ref["cells"])
    assert ref["n_cells"].mean() > 5


@pytest.mark.parametrize("kind", ["pinhole", "fisheye"])
def test_trace_forward_matches_fp64_oracle(kind):
    from powerfoam_amd.rasterize import power_face_encoding
    s = big_scene(2)
    cam = cameras(kind)
    x = _tensors(s, "cuda")
    diff = power_face_encoding(x["points"], x["radii"], x["adj"], x["off"])
    st = start_cell(s, cam.eye.numpy().astype(np.float64))
    tr = _tracer(8)
    col, opac, depth, normal = tr.forward(_gpu_cam(cam), st, x["points"], x["radii"], x["density"], x["normals"],
                                          x["texel_sites"], x["texel_rgb"], x["texel_height"], x["adj"], x["off"],
                                          diff, 1e-3, 0.5)
    bench = tr.benchmark(_gpu_cam(cam), st, x["points"], x["radii"], x["density"], x["normals"], x["texel_sites"],
                         x["texel_rgb"], x["texel_height"], x["adj"], x["off"], diff, 1e-3)
    ocam, _k = _oracle_cam(cam)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    ns = np.concatenate([s["normals"], s["density"][:, None]], 1)
    rgbh = np.concatenate([s["texel_rgb"], s["texel_height"][..., None]], -1)
    inp = orc.Inputs(sp, ns, s["texel_sites"], rgbh, s["adj"], s["off"], diff.cpu().numpy().view(np.uint16))
    ref = orc.trace_fwd(ocam, inp, st, threshold=1e-3, depth_q=0.5)
    assert np.array_equal(bench.cpu().numpy(), col.detach().cpu().numpy())
    assert np.abs(col.detach().cpu().numpy() - ref["color"]).max() < ABS_TOL
    assert np.abs(opac.detach().cpu().numpy() - (1 - np.exp(ref["log_t"]))).max() < ABS_TOL
    assert np.abs(normal.detach().cpu().numpy() - ref["normal"]).max() < ABS_TOL
    d = depth.detach().cpu().numpy()
    assert np.array_equal(d == -1, ref["depth"] == -1)
    assert np.abs(d - ref["depth"]).max() < ABS_TOL
    assert (ref["depth"] != -1).mean() > 0.1


@pytest.mark.parametrize("kind", ["pinhole", "fisheye"])
def test_trace_backward_matches_fp64_adjoint(kind):
    import torch
    s = big_scene(3)
    cam = cameras(kind)
    x = _tensors(s, "cuda")
    leaves = [x[k].requires_grad_(True) for k in ("points", "radii", "density", "normals", "texel_sites",
                                                 "texel_rgb", "texel_height")]
    st = start_cell(s, cam.eye.numpy().astype(np.float64))
    col, opac, depth, normal = _tracer(8).forward(_gpu_cam(cam), st, *leaves, x["adj"], x["off"], None, 1e-3, 0.5)
    H, W = cam.height, cam.width
    ocam, _k = _oracle_cam(cam)
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1)
    ns = np.concatenate([s["normals"], s["density"][:, None]], 1)
    rgbh = np.concatenate([s["texel_rgb"], s["texel_height"][..., None]], -1)
    inp = orc.Inputs(sp, ns, s["texel_sites"], rgbh, s["adj"], s["off"], None)
    ref = orc.trace_fwd(ocam, inp, st, threshold=1e-3, depth_q=0.5)
    # rays through ill-conditioned cells (grazing balls, tied bounds) carry no
    # loss weight on either side
    keep = ~ref["fragile"]
    assert keep.mean() > 0.5
    rng = np.random.default_rng(7)
    w = dict(color=rng.normal(size=(H, W, 3)) * keep[..., None], normal=rng.normal(size=(H, W, 3)) * keep[..., None],
             opacity=rng.normal(size=(H, W)) * keep, depth=rng.normal(size=(H, W)) * keep)
    valid = (depth.detach() != -1)
    tw = {k: torch.tensor(v, dtype=torch.float32, device="cuda") for k, v in w.items()}
    L = (col * tw["color"]).sum() + (normal * tw["normal"]).sum() + (opac * tw["opacity"]).sum() + \
        (torch.where(valid, depth, torch.zeros_like(depth)) * tw["depth"]).sum()
    L.backward()
    wd = np.where(ref["depth"] != -1, w["depth"], 0.0)
    g = orc.trace_bwd(ocam, inp, st, ref, w["color"], w["normal"], wd, w["opacity"], threshold=1e-3, depth_q=0.5)
    got = dict(grad_spheres=np.concatenate([leaves[0].grad.cpu().numpy(), leaves[1].grad.cpu().numpy()[:, None]], 1),
               grad_nsigmas=np.concatenate([leaves[3].grad.cpu().numpy(), leaves[2].grad.cpu().numpy()[:, None]], 1),
               grad_texel_sites=leaves[4].grad.cpu().numpy(), grad_texel_rgb=leaves[5].grad.cpu().numpy(),
               grad_texel_height=leaves[6].grad.cpu().numpy())
    for k in harness.GRAD_KEYS:
        err = harness.rel_err(got[k], g[k])
        print(k, err, flush=True)
        assert err < GRAD_REL_TOL, (k, err)


@pytest.mark.parametrize("kind", ["pinhole", "fisheye"])
@pytest.mark.parametrize("encoded", [True, False], ids=["fp16faces", "planes"])
def test_compacted_walk_equals_per_tile_walk(kind, encoded):
    """Ballot compaction (a wave walks 4 tiles' rays and refills ended lanes,
    skipping slots outside the image)
    changes which lane walks a ray, not the ray: every output and the visited
    cells are bit-identical to one wave per tile, with the fp16 face records
    and with the exact face planes."""
    from powerfoam_amd.rasterize import power_face_encoding
    s = big_scene(4)
    cam = cameras(kind)
    x = _tensors(s, "cuda")
    diff = power_face_encoding(x["points"], x["radii"], x["adj"], x["off"]) if encoded else None
    st = start_cell(s, cam.eye.numpy().astype(np.float64))
    res = {}
    for compact in (4, 0):
        tr = _tracer(8)
        tr.compact = compact
        print(f"walk {kind} encoded={encoded} compact={compact}", flush=True)
        outs = tr.forward(_gpu_cam(cam), st, x["points"], x["radii"], x["density"], x["normals"], x["texel_sites"],
                          x["texel_rgb"], x["texel_height"], x["adj"], x["off"], diff, 1e-3, 0.5)
        cells, n = tr.trace_cells(_gpu_cam(cam), st, x["points"], x["radii"], x["density"], x["normals"],
                                  x["texel_sites"], x["texel_rgb"], x["texel_height"], x["adj"], x["off"], diff, 1e-3,
                                  max_cells=96)
        res[compact] = [o.detach().cpu().numpy() for o in outs] + [cells.cpu().numpy(), n.cpu().numpy()]
    for a, b in zip(res[4], res[0]):
        assert np.array_equal(a, b)
    assert res[4][-1].mean() > 3


@pytest.mark.parametrize("kind", ["pinhole_ragged", "fisheye"])
def test_tile_steps_and_longest_first_backward(kind):
    """The forward's per-tile longest walk equals the maximum of the visited
    cell counts over the tile's pixels (column-major 8x8 tiles, a ragged image
    edge included); pf_trace_tile_order returns a permutation by descending
    walk length; the backward taken in that order gives the gradients of the
    index-order backward up to summation order, and the forward is unchanged."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd.rasterize import _ref
    s = big_scene(5)
    if kind == "pinhole_ragged":
        cam = look_at((0.1, 0.2, 2.2), (0.0, 0.0, 0.0), 70, 51, 60.0)
    else:
        cam = cameras("fisheye")
    x = _tensors(s, "cuda")
    st = start_cell(s, cam.eye.numpy().astype(np.float64))
    gcam = _gpu_cam(cam)
    H, W = cam.height, cam.width
    th, tw = -(-H // 8), -(-W // 8)
    tr = _tracer(8)
    lib = _abi.load()
    print(f"tile steps {kind}", flush=True)
    t = tr._prep(x["points"], x["radii"], x["density"], x["normals"], x["texel_sites"], x["texel_rgb"],
                 x["texel_height"], x["adj"], x["off"], None)
    a, keep = tr._args(gcam, st, t, 1e-3, 0.5)
    color = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    n_cells = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    steps = torch.full((th * tw,), -7, dtype=torch.int32, device="cuda")
    a.color, a.n_cells, a.tile_steps = color.data_ptr(), n_cells.data_ptr(), steps.data_ptr()
    _abi.check(lib.pf_trace_fwd(_ref(a), _abi.stream_handle()))
    n = n_cells.cpu().numpy()
    pad = np.full((th * 8, tw * 8), -1, np.int64)
    pad[:H, :W] = n
    want = pad.reshape(th, 8, tw, 8).max(axis=(1, 3)).T.reshape(-1)  # tile t = row t % th, column t // th
    got = steps.cpu().numpy()
    assert np.array_equal(got, want)
    order = torch.empty_like(steps)
    _abi.check(lib.pf_trace_tile_order(steps.data_ptr(), steps.numel(), order.data_ptr(), _abi.stream_handle()))
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(th * tw))
    assert np.all(np.diff(np.minimum(got[o], 1023)) <= 0)

    # backward in that order vs index order
    res = {}
    for lf in (True, False):
        print(f"backward {kind} longest_first={lf}", flush=True)
        tr = _tracer(8)
        tr.longest_first = lf
        leaves = [x[k].detach().clone().requires_grad_(True) for k in ("points", "radii", "density", "normals",
                                                                      "texel_sites", "texel_rgb", "texel_height")]
        outs = tr.forward(gcam, st, *leaves, x["adj"], x["off"], None, 1e-3, 0.5)
        rng = np.random.default_rng(11)
        L = sum((o * torch.tensor(rng.normal(size=tuple(o.shape)), dtype=torch.float32, device="cuda")).sum()
                for o in outs if o is not outs[2])
        L.backward()
        res[lf] = ([o.detach().cpu().numpy() for o in outs], [v.grad.cpu().numpy() for v in leaves])
    for a_, b_ in zip(res[True][0], res[False][0]):
        assert np.array_equal(a_, b_)
    for ga, gb in zip(res[True][1], res[False][1]):
        assert np.allclose(ga, gb, rtol=1e-4, atol=1e-5 * max(np.abs(gb).max(), 1e-30))


@pytest.mark.parametrize("kind", ["pinhole", "fisheye"])
def test_face_planes_walk_equals_sphere_walk(kind):
    """pf_trace_planes writes, per directed face, n = c_a - c_s and
    pm = (|c_a|^2 - r_a^2)/2 - (|c_s|^2 - r_s^2)/2 bit-exactly as float32
    numpy forms them in the same order; a walk that reads those planes
    (RayTracer.face_planes) visits the same cells, bit for bit, as the walk
    that forms them from the spheres, and returns the same outputs and
    gradients to float rounding (the shading outside the walk may contract
    multiply-adds differently in the two kernel variants)."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd.rasterize import _ref, _scene_struct
    s = big_scene(6)
    cam = cameras(kind)
    x = _tensors(s, "cuda")
    st = start_cell(s, cam.eye.numpy().astype(np.float64))
    gcam = _gpu_cam(cam)
    tr = _tracer(8)
    t = tr._prep(x["points"], x["radii"], x["density"], x["normals"], x["texel_sites"], x["texel_rgb"],
                 x["texel_height"], x["adj"], x["off"], None)
    planes = t[7].cpu().numpy()
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1).astype(np.float32)
    owner = np.repeat(np.arange(len(sp)), np.diff(s["off"]))
    c, a = sp[owner], sp[s["adj"]]
    half = np.float32(0.5)
    pm = half * ((a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1] + a[:, 2] * a[:, 2]) - a[:, 3] * a[:, 3]) - \
        half * ((c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1] + c[:, 2] * c[:, 2]) - c[:, 3] * c[:, 3])
    want = np.concatenate([a[:, :3] - c[:, :3], pm[:, None]], 1)
    assert np.array_equal(planes.view(np.uint32), want.astype(np.float32).view(np.uint32))

    res = {}
    for fp in (True, False):
        print(f"walk {kind} face_planes={fp}", flush=True)
        tr = _tracer(8)
        tr.face_planes = fp
        leaves = [x[k].detach().clone().requires_grad_(True) for k in ("points", "radii", "density", "normals",
                                                                      "texel_sites", "texel_rgb", "texel_height")]
        outs = tr.forward(gcam, st, *leaves, x["adj"], x["off"], None, 1e-3, 0.5)
        rng = np.random.default_rng(13)
        L = sum((o * torch.tensor(rng.normal(size=tuple(o.shape)), dtype=torch.float32, device="cuda")).sum()
                for o in (outs[0], outs[1], outs[3]))
        L.backward()
        cells, n = tr.trace_cells(gcam, st, *[v.detach() for v in leaves], x["adj"], x["off"], None, 1e-3,
                                  max_cells=96)
        res[fp] = ([o.detach().cpu().numpy() for o in outs] + [cells.cpu().numpy(), n.cpu().numpy()],
                   [v.grad.cpu().numpy() for v in leaves])
    for a_, b_ in zip(res[True][0][4:], res[False][0][4:]):  # visited cells and their count
        assert np.array_equal(a_, b_)
    for a_, b_ in zip(res[True][0][:4], res[False][0][:4]):  # colour, opacity, depth, normals
        assert np.allclose(a_, b_, rtol=1e-5, atol=1e-6)
    for ga, gb in zip(res[True][1], res[False][1]):
        assert np.allclose(ga, gb, rtol=1e-4, atol=1e-5 * max(np.abs(gb).max(), 1e-30))
    assert res[True][0][-1].mean() > 3


def test_face_planes_stop_at_the_face_count():
    """pf_trace_planes follows no offset past n_faces: with CSR offsets that
    claim more faces than the adjacency holds (the last site's range runs
    past its end), the planes of the faces that exist are written as with a
    consistent CSR and the guard words after them stay untouched. Every read
    stays inside the arrays (ids below n_faces, offsets of real sites)."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd.rasterize import _ref, _scene_struct
    s = big_scene(7, n=500)
    x = _tensors(s, "cuda")
    tr = _tracer(8)
    t = tr._prep(x["points"], x["radii"], x["density"], x["normals"], x["texel_sites"], x["texel_rgb"],
                 x["texel_height"], x["adj"], x["off"], None)
    want = t[7].clone()
    E = int(x["adj"].numel())
    bad_off = x["off"].clone()
    bad_off[-1] = E + 100  # the last site claims 100 faces that do not exist
    lib = _abi.load()
    sc = _scene_struct(t[0], t[1], t[2], t[3], t[4], bad_off.contiguous(), None)
    guard = 64
    planes = torch.full((E + guard, 4), float("nan"), dtype=torch.float32, device="cuda")
    _abi.check(lib.pf_trace_planes(_ref(sc), E, planes.data_ptr(), _abi.stream_handle()))
    torch.cuda.synchronize()
    got = planes.cpu().numpy()
    assert np.array_equal(got[:E].view(np.uint32), want.cpu().numpy().view(np.uint32))
    assert np.isnan(got[E:]).all()
