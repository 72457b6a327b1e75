"""Benchmark of the Power Foam rasterizer hot path on one MI355X.

Workload (one "step" = one frame): the 1M-site scene (seed 1; 90% from a
64-component Gaussian mixture, 10% on a radius 6-10 shell; r log-uniform
[0.002, 0.08]; density lognormal(0, 1.5); degree-3 SH colour), 32 pinhole
cameras on a radius-4 circle at height 1, 1297x840, fov 70 deg. Per frame:
SH colour head (texel_rgb for this camera) -> Rasterizer.forward (binning,
one host sync for the list size, adjacency encoding, compositing) -> L2 loss
against a uniform-random target -> loss.backward() (compositing adjoint, SH
adjoint). All scene tensors are resident in HBM before the timed region.

Prints one JSON line (rank 0). `value` is rays per second (Mrays/s) over the
timed frames; `ms_per_step` is ms per frame. The roofline entry is for the
kernel that takes the most time: algorithmic bytes per launch from
`algorithmic_bytes` (per-pair, per-pixel and per-site byte counts applied to
the measured workload counts of the views the timed frames rendered),
divided by that kernel's average duration measured with HIP events on its
launch stream; beside it the VALU issue fraction (wave-instructions per
launch from a committed rocprofv3 counter run of the same kernels and views,
over a VALU peak the run measures itself with an FMA kernel). The default
line also carries the learned-head record (3M-site scene, MLP head trained
on the GPU, fwd+bwd frames with each head, render PSNR against the SH
render, MFMA utilisation of the head's kernels) and the ray-tracer record
(the same 1M-site scene with its full power adjacency, 16 equidistant
180-degree fisheye views of 1024x1024 rays, fwd+bwd; `trace_record`). The
CPU baseline is the FP32 CPU restatement of the same algorithm (the oracle,
OpenMP) on a bounded sample of the same workload.
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "fwd+bwd ms/frame @1M sites 1297x840; Mrays/s; MLP vs SH head at PSNR≥45dB"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16
LATENT, HIDDEN = 16, 64


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(n_sites, n_views, width, height, dev, seed=1):
    import torch
    from powerfoam_amd import synth

    t0 = time.time()
    s = synth.scene_gmm(seed, n=n_sites)
    log(f"scene: {n_sites} sites in {time.time() - t0:.1f}s")
    t0 = time.time()
    adj, off, stats = synth.power_adjacency(s["points"], s["radii"], return_stats=True)
    log(f"power/alpha adjacency: {stats} in {time.time() - t0:.1f}s")
    cams = synth.circle_cameras(n_views=n_views, width=width, img_h=height)
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    rng = np.random.default_rng(1234)
    target = tt(rng.uniform(0.0, 1.0, (height, width, 3)).astype(np.float32))
    params = dict(points=tt(s["points"]), radii=tt(s["radii"]), density=tt(s["density"]),
                  normals=tt(s["normals"]), texel_sites=tt(s["texel_sites"]), sh=tt(s["sh"]),
                  texel_height=tt(s["texel_height"]))
    for v in params.values():
        v.requires_grad_(True)
    return dict(scene=s, params=params, adjacency=tt(adj, torch.int32), adjacency_offsets=tt(off, torch.int32),
                adj_np=adj, off_np=off, cams=cams, target=target, stats=stats)


def frame(rast, head, w, cam):
    """One training-style frame: colour head, rasterize, L2 loss, backward.
    Gradients start from None, as after an optimiser step with
    zero_grad(set_to_none=True)."""
    import torch
    p = w["params"]
    for v in p.values():
        v.grad = None
    if hasattr(head, "zero_grad"):
        head.zero_grad()
    if getattr(head, "fused", False):
        # the learned head evaluated inside the compositing forward
        out = rast.forward(cam, None, p["points"], p["radii"], p["density"], p["normals"], p["texel_sites"],
                           None, p["texel_height"], w["adjacency"], w["adjacency_offsets"], w["target"], True,
                           color_head=head.mlp)
    else:
        texel_rgb = head.forward(p["texel_sites"], cam, p["sh"])
        out = rast.forward(cam, None, p["points"], p["radii"], p["density"], p["normals"], p["texel_sites"],
                           texel_rgb, p["texel_height"], w["adjacency"], w["adjacency_offsets"], w["target"], True)
    loss = torch.mean((out[0] - w["target"]) ** 2)
    loss.backward()
    return out


class MlpSlot:
    """The learned head in the colour-head slot of `frame` (same call shape as SHColor)."""

    def __init__(self, mlp):
        self.mlp = mlp

    def forward(self, texel_sites, camera, sh):
        return self.mlp.forward(texel_sites, camera)

    def zero_grad(self):
        for q in self.mlp.parameters():
            q.grad = None


class MlpFused(MlpSlot):
    """The learned head inside the rasterizer (color_head=): in the
    compositing forward or over the binned sites, per Rasterizer.head_mode."""
    fused = True


def mlp_bwd_rows(rast, mlp, w, cams):
    """Mean over `cams` of the texel rows pf_mlp_bwd processes (every texel of
    the sites with a non-zero colour gradient: pf_mlp_active_sites) and of the
    rows with a non-zero gradient, from one pre-pass frame per view."""
    import torch
    p = w["params"]
    done, nonzero = [], []
    for cam in cams:
        rgb = mlp.forward(p["texel_sites"], cam)
        rgb.retain_grad()
        out = rast.forward(cam, None, p["points"], p["radii"], p["density"], p["normals"], p["texel_sites"], rgb,
                           p["texel_height"], w["adjacency"], w["adjacency_offsets"], w["target"], True)
        torch.mean((out[0] - w["target"]) ** 2).backward()
        g = rgb.grad != 0
        T = g.shape[1]
        done.append(int(g.any(dim=2).any(dim=1).sum()) * T)
        nonzero.append(int(g.any(dim=2).sum()))
        for q in list(mlp.parameters()) + list(p.values()):
            q.grad = None
    return float(np.mean(done)), float(np.mean(nonzero))


def time_frames(rast, head, w, steps, heads_log=None):
    """Wall time of `steps` frames (synchronised on both sides) and the mean
    kernel times the rasterizer (and an MLP head) recorded with HIP events."""
    import torch
    cams = w["cams"]
    rast.event_log = []
    if heads_log is not None:
        heads_log.event_log = rast.event_log
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        frame(rast, head, w, cams[i % len(cams)])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ev = rast.event_log
    rast.event_log = None
    if heads_log is not None:
        heads_log.event_log = None
    kern = {}
    for name, s, e in ev:
        kern.setdefault(name, []).append(s.elapsed_time(e))
    return elapsed, {k: float(np.mean(v)) for k, v in kern.items()}


def render_psnr(rast, w, head_a, head_b, cams=None):
    """Mean over views of the PSNR of head_a's render against head_b's (forward only)."""
    import torch
    from powerfoam_amd.mlp_head import psnr
    p = w["params"]
    vals = []
    with torch.no_grad():
        for cam in (cams if cams is not None else w["cams"]):
            imgs = []
            for h in (head_a, head_b):
                rgb = h.forward(p["texel_sites"], cam, p["sh"]).detach()
                out = rast.forward(cam, None, p["points"], p["radii"], p["density"], p["normals"], p["texel_sites"],
                                   rgb, p["texel_height"], w["adjacency"], w["adjacency_offsets"], w["target"], False)
                imgs.append(out[0])
            vals.append(psnr(imgs[0], imgs[1]))
    return float(np.mean(vals))


def mlp_flops_per_row():
    """Algorithmic FLOPs per texel row of the learned head: forward
    2 (64*32 + 64*64 + 3*64); backward the recomputed forward plus the input
    gradients of layers 3 and 2 and of the latent slots of layer 1, and the
    three weight gradients."""
    fwd = 2 * (HIDDEN * 2 * LATENT + HIDDEN * HIDDEN + 3 * HIDDEN)
    bwd = fwd + 2 * (3 * HIDDEN + HIDDEN * HIDDEN + HIDDEN * LATENT) + \
        2 * (3 * HIDDEN + HIDDEN * HIDDEN + HIDDEN * 2 * LATENT)
    return fwd, bwd


def mlp_head_record(args, dev, psnr_views=4):
    """Learned colour head on the 3M-site scene (seed 2): train on SH-labelled
    (site, direction) pairs with the head's own kernels, render PSNR of the
    MLP head (and of the view-independent DC colour) against the SH render,
    then fwd+bwd frames with each head. Returns (record, ms per MLP frame)."""
    import torch
    from types import SimpleNamespace
    from powerfoam_amd.color_fn import SHColor
    from powerfoam_amd.mlp_head import MLPColor, fit, latents_from_sh
    from powerfoam_amd.rasterize import Rasterizer

    w = build_workload(args.mlp_sites, args.views, args.width, args.height, dev, seed=2)
    rast = Rasterizer(SimpleNamespace(num_texel_sites=8, render_objective="volume", is_pinhole=True), dev)
    sh = w["params"]["sh"].detach()
    sh_head = SHColor()
    cams = w["cams"][:psnr_views]
    # each direction encoding trained the same way; the frames are timed with
    # args.mlp_encoding, the other one's PSNR is reported beside it
    psnr_by_enc, trained = {}, {}
    for enc in [args.mlp_encoding] + [e for e in ("freq", "sh") if e != args.mlp_encoding]:
        m = MLPColor(sh.shape[0], dev, seed=0, latents=latents_from_sh(sh), encoding=enc)
        t0 = time.time()
        h = fit(m, sh, steps=args.mlp_steps, batch=args.mlp_batch, log_every=max(1, args.mlp_steps // 8))
        torch.cuda.synchronize()
        trained[enc] = (m, h, time.time() - t0)
        psnr_by_enc[enc] = render_psnr(rast, w, MlpSlot(m), sh_head, cams)
        log(f"trained the MLP head ({enc} encoding): {args.mlp_steps} steps x {args.mlp_batch} pairs in "
            f"{trained[enc][2]:.1f}s, loss {h}, render PSNR {psnr_by_enc[enc]:.2f} dB")
        if enc != args.mlp_encoding:
            del m
    mlp, hist, train_s = trained[args.mlp_encoding]
    mlp_head = MlpSlot(mlp)

    class _DC:
        def forward(self, texel_sites, camera, sh_):
            n, T = texel_sites.shape[:2]
            return torch.clamp(sh_[:, 0, :].detach(), min=0.0)[:, None, :].expand(n, T, 3).contiguous()

    psnr_mlp = psnr_by_enc[args.mlp_encoding]
    psnr_dc = render_psnr(rast, w, _DC(), sh_head, cams)
    log(f"render PSNR vs the SH render ({len(cams)} views): MLP {psnr_mlp:.2f} dB, view-independent DC "
        f"{psnr_dc:.2f} dB")
    steps = args.mlp_frames
    fused_head = MlpFused(mlp)
    for i in range(args.warmup):
        frame(rast, sh_head, w, w["cams"][i % len(w["cams"])])
        frame(rast, mlp_head, w, w["cams"][i % len(w["cams"])])
        for mode in ("compositing", "binned"):
            rast.head_mode = mode
            frame(rast, fused_head, w, w["cams"][i % len(w["cams"])])
    sh_s, sh_k = time_frames(rast, sh_head, w, steps)
    mlp_s, mlp_k = time_frames(rast, mlp_head, w, steps, heads_log=mlp)
    # the head inside the rasterizer: in the compositing forward, then once
    # per site with a tile between binning and compositing
    rast.head_mode = "compositing"
    fused_s, fused_k = time_frames(rast, fused_head, w, steps, heads_log=mlp)
    rast.head_mode = "binned"
    binned_s, binned_k = time_frames(rast, fused_head, w, steps, heads_log=mlp)
    # the backward skips 32-row steps without a colour gradient: its MFMA work
    # is over the rows of the steps it processes (views of the timed frames)
    bwd_rows, nz_rows = mlp_bwd_rows(rast, mlp, w, [w["cams"][i % len(w["cams"])] for i in range(min(steps, 8))])
    rows = sh.shape[0] * 8
    f_fwd, f_bwd = mlp_flops_per_row()
    mfma = {}
    for name, f, n_rows in (("mlp_fwd", f_fwd, rows), ("mlp_bwd", f_bwd, bwd_rows)):
        if name in mlp_k:
            ach = f * n_rows / (mlp_k[name] * 1e-3) / 1e12
            mfma[name] = dict(rows=n_rows, flops_per_launch=f * n_rows, ms=mlp_k[name], achieved=ach,
                              peak=MFMA_BF16_PEAK_TFLOPS, unit="TFLOP/s", frac=ach / MFMA_BF16_PEAK_TFLOPS)
    rec = dict(workload=f"{args.mlp_sites}-site GMM scene (seed 2), {args.views} pinhole views "
                        f"{args.width}x{args.height}, fwd+bwd per frame with each colour head",
               mlp_ms_per_frame=mlp_s * 1e3 / steps, sh_ms_per_frame=sh_s * 1e3 / steps,
               speedup_vs_sh_head=sh_s / mlp_s, frames=steps,
               mlp_fused_ms_per_frame=fused_s * 1e3 / steps, fused_speedup_vs_sh_head=sh_s / fused_s,
               kernel_ms_mlp_fused=fused_k,
               mlp_binned_ms_per_frame=binned_s * 1e3 / steps, binned_speedup_vs_sh_head=sh_s / binned_s,
               kernel_ms_mlp_binned=binned_k, mlp_bwd_rows_nonzero=nz_rows, mlp_bwd_rows_processed=bwd_rows,
               texel_rows=rows,
               psnr_vs_sh_render_db=psnr_mlp, psnr_dc_only_vs_sh_render_db=psnr_dc, psnr_views=len(cams),
               encoding=args.mlp_encoding, psnr_vs_sh_render_db_by_encoding=psnr_by_enc,
               training_pairs=args.mlp_steps * args.mlp_batch, train_seconds=train_s,
               train_loss_first_last=[hist[0][1], hist[-1][1]] if hist else None,
               kernel_ms_mlp=mlp_k, kernel_ms_sh=sh_k, mfma=mfma,
               head="16-dim bf16 latent + 16-feature direction encoding (freq: 1, d, sin/cos(pi d), "
                    "sin/cos(2 pi d); sh: 1, Y_1..Y_15) -> 64 -> 64 -> 3, bf16 MFMA, fp32 accumulate")
    log(f"learned head: MLP {rec['mlp_ms_per_frame']:.3f} ms (fused into compositing "
        f"{rec['mlp_fused_ms_per_frame']:.3f} ms, over the binned sites {rec['mlp_binned_ms_per_frame']:.3f} ms) "
        f"vs SH {rec['sh_ms_per_frame']:.3f} ms per frame; "
        f"MFMA {({k: round(v['frac'], 4) for k, v in mfma.items()})}; kernels pre-pass {mlp_k}, fused {fused_k}, "
        f"SH {sh_k}; backward rows {bwd_rows:.0f} processed, {nz_rows:.0f} non-zero of {rows}")
    del w, rast, mlp
    torch.cuda.empty_cache()
    return rec, mlp_s * 1e3 / steps


def run_trace(args, dev):
    """`--workload trace`: the ray tracer workload below as the bench line."""
    line = trace_record(args, dev, args.steps, args.warmup)
    print(json.dumps(line), flush=True)


def trace_record(args, dev, steps, warmup):
    """Ray tracer workload: the 1M-site GMM scene (seed 1) with its full power
    adjacency (the tracer walks power cells; alpha filtering would cut cells
    open), one equidistant 180-degree fisheye view of 1024x1024 rays per camera
    of the circle, starting in the cell of the eye (benchmark.py:320-324), with
    exact neighbour geometry (the fp16 face encoding of RayTracer.benchmark
    rounds the faces around a cell edge independently, so they no longer meet
    in one line, and on this scene a few rays circle such an edge until the
    walk cap); per view the SH head, the tracer forward (colour, opacity,
    depth, normals) and the backward of an L2 colour loss. Reports rays/s over
    the timed views and the mean cells visited per ray. Algorithmic bytes
    count, per distinct cell a wave of 64 rays visits (the kernels walk an 8x8
    block of rays together): site record 32 B + CSR offsets 8 B + 20 B per face
    (neighbour id and sphere) + 8 texels 224 B, and 40 B of outputs per ray
    (forward); the backward reads as much and read-modify-writes the 256 B of
    site and texel gradients per distinct cell, with 64 B of saved and
    incoming data per ray."""
    import torch
    from types import SimpleNamespace
    from powerfoam_amd import synth
    from powerfoam_amd.camera import TorchCamera, fisheye_ray_maps, look_at
    from powerfoam_amd.color_fn import SHColor
    from powerfoam_amd.raytrace import RayTracer, start_cell

    size = 1024
    t0 = time.time()
    s = synth.scene_gmm(1, n=args.sites)
    adj, off, stats = synth.power_adjacency(s["points"], s["radii"], alpha_complex=False, return_stats=True)
    log(f"scene + full power adjacency {stats} in {time.time() - t0:.1f}s")
    tt = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    p = dict(points=tt(s["points"]), radii=tt(s["radii"]), density=tt(s["density"]), normals=tt(s["normals"]),
             texel_sites=tt(s["texel_sites"]), sh=tt(s["sh"]), texel_height=tt(s["texel_height"]))
    for v in p.values():
        v.requires_grad_(True)
    adjacency, offsets = tt(adj, torch.int32), tt(off, torch.int32)
    cams, starts = [], []
    for c in synth.circle_cameras(n_views=args.views):
        eye = c.eye.numpy().astype(np.float64)
        rm = torch.tensor(fisheye_ray_maps(eye, (0.0, 0.0, 0.0), size), device=dev)
        t = look_at(eye, (0.0, 0.0, 0.0), size, size, 90.0)
        cams.append(TorchCamera(t.eye, t.right, t.up, size, size, rm))
        starts.append(start_cell(p["points"].detach(), p["radii"].detach(), eye))
    tracer = RayTracer(SimpleNamespace(num_texel_sites=8, is_pinhole=True), dev)
    head = SHColor()
    target = tt(np.random.default_rng(1234).uniform(0.0, 1.0, (size, size, 3)).astype(np.float32))

    def view(i):
        cam = cams[i % len(cams)]
        rgb = head.forward(p["texel_sites"], cam, p["sh"])
        col, opac, depth, normal = tracer.forward(cam, starts[i % len(cams)], p["points"], p["radii"], p["density"],
                                                  p["normals"], p["texel_sites"], rgb, p["texel_height"], adjacency,
                                                  offsets, None, 1e-3, 0.5)
        torch.mean((col - target) ** 2).backward()

    for i in range(warmup):
        view(i)
    torch.cuda.synchronize()
    tracer.event_log = []
    t0 = time.perf_counter()
    for i in range(steps):
        view(i)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kern = {}
    for name, a, b in tracer.event_log:
        kern.setdefault(name, []).append(a.elapsed_time(b))
    tracer.event_log = None
    kern_ms = {k: float(np.mean(v)) for k, v in kern.items()}
    # visited cells per ray, and the cells each wave (an 8x8 block of rays,
    # which the kernels run together) visits at least once: the bytes a wave
    # must fetch once, over the timed views
    deg = (offsets[1:] - offsets[:-1]).to(torch.int64)
    visits, uniq, uniq_faces = [], [], []
    K = 256
    with torch.no_grad():
        for i in range(min(steps, len(cams))):
            cam = cams[i]
            rgb = head.forward(p["texel_sites"], cam, p["sh"])
            cells, n_cells = tracer.trace_cells(cam, starts[i], p["points"], p["radii"], p["density"], p["normals"],
                                                p["texel_sites"], rgb, p["texel_height"], adjacency, offsets, None,
                                                1e-3, max_cells=K)
            visits.append(float(n_cells.to(torch.float64).sum()))
            w = cells.view(size // 8, 8, size // 8, 8, K).permute(2, 0, 1, 3, 4).reshape(-1, 64 * K)
            w = torch.sort(w, dim=1).values
            first = torch.ones_like(w, dtype=torch.bool)
            first[:, 1:] = w[:, 1:] != w[:, :-1]
            first &= w >= 0
            u = w[first].to(torch.int64)
            uniq.append(float(u.numel()))
            uniq_faces.append(float(deg[u].sum()))
            del cells, w, first
    rays = size * size
    v_mean, u_mean, uf_mean = float(np.mean(visits)), float(np.mean(uniq)), float(np.mean(uniq_faces))
    fwd_bytes = u_mean * (32 + 8 + 224) + 20 * uf_mean + rays * 40
    bwd_bytes = u_mean * (32 + 8 + 224 + 2 * 256) + 20 * uf_mean + rays * 64
    abytes = dict(trace_fwd=fwd_bytes, trace_bwd=bwd_bytes)
    dom = max(kern_ms, key=kern_ms.get)
    achieved = abytes[dom] / (kern_ms[dom] * 1e-3) / 1e9
    ms = elapsed * 1e3 / steps
    line = dict(metric="ray tracer fwd+bwd, Mrays/s", value=rays * steps / elapsed / 1e6, unit="Mrays/s",
                n_gpus=1, steps=steps, warmup=warmup, ms_per_step=ms, higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f32", data="synthetic (seeded scene generator)",
                config=dict(workload=f"ray tracer fwd+bwd with SH-3 colour head, {args.sites}-site GMM scene, "
                                     f"{args.views} equidistant 180-degree fisheye views {size}x{size}",
                            sites=args.sites, rays=rays, views=args.views, adjacency="full power",
                            adjacency_edges=int(stats["edges"]), cells_per_ray=v_mean / rays,
                            distinct_cells_per_wave=u_mean / (rays / 64),
                            faces_per_distinct_cell=uf_mean / max(u_mean, 1.0)),
                kernel_ms=kern_ms,
                roofline=dict(bound="hbm", kernel=dom, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=achieved / HBM_PEAK_GBS, traffic=None, algorithmic_bytes=abytes[dom]))
    log(f"{ms:.3f} ms/view, {line['value']:.1f} Mrays/s; kernels {kern_ms}; cells/ray {v_mean / rays:.1f}")
    return line


def run_cech(args, dev):
    """Cech adjacency build (AABBTree.build_cech_complex, bvh.py:96-123) of the
    GMM scene (seed 1, --sites sites): count pass, one host sync for the
    total, emit pass, per-site ascending sort. Reports adjacency entries/s.
    Algorithmic bytes per build: the spheres read once (16 B per site), the
    offsets and lists written once (4 B per site, 4 B per entry). The CPU
    baseline is the oracle's kd-tree query of a 4,000-site sample."""
    import torch
    from powerfoam_amd import synth
    from powerfoam_amd.bvh import AABBTree

    s = synth.scene_gmm(1, n=args.sites)
    tree = AABBTree(dev)
    tree.update(torch.tensor(s["points"], device=dev), torch.tensor(s["radii"], device=dev))
    for _ in range(args.warmup):
        adj, off = tree.build_cech_complex()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        adj, off = tree.build_cech_complex()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    m = adj.numel()
    abytes = 20 * args.sites + 4 * m
    achieved = abytes / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import cech as oc
        sample = np.random.default_rng(0).choice(args.sites, 4000, replace=False)
        t0 = time.perf_counter()
        got = oc.neighbours(s["points"], s["radii"], sample)
        dt = time.perf_counter() - t0
        cpu = dict(value=sum(len(g) for g in got) / dt / 1e6, unit="M entries/s", cores=1, kind="port",
                   sample="oracle kd-tree neighbour lists of 4,000 random sites of the same scene")
    line = dict(metric="Cech adjacency build, M entries/s", value=m / (ms * 1e-3) / 1e6, unit="M entries/s",
                n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=ms, higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f32", data="synthetic (seeded scene generator)",
                config=dict(workload=f"Cech adjacency of the {args.sites}-site GMM scene", sites=args.sites,
                            entries=m, mean_degree=m / args.sites),
                roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=achieved / HBM_PEAK_GBS, traffic=None, algorithmic_bytes=abytes,
                              bound_note="whole build; the neighbour search (candidate cells through L2) sets "
                                         "the time, not HBM"),
                cpu_baseline=cpu)
    log(f"{ms:.2f} ms per build, {m} entries")
    print(json.dumps(line), flush=True)


def workload_counts(rast, w, width, height):
    """Per-frame counts for the algorithmic byte formulas, from the last frame."""
    import torch
    offsets, sites, early = rast.last_tiles
    counts = (offsets[1:] - offsets[:-1]).to(torch.int64)
    tile_of = torch.repeat_interleave(torch.arange(counts.numel(), device=counts.device), counts)
    pos = torch.arange(sites.numel(), device=sites.device) - offsets[:-1][tile_of]
    processed = sites[pos < early.to(torch.int64)[tile_of]].to(torch.int64)
    off = w["adjacency_offsets"].to(torch.int64)
    nadj = (off[processed + 1] - off[processed])
    touched = torch.unique(processed)
    deg = off[touched + 1] - off[touched]
    nb = w["adjacency"][torch.repeat_interleave(off[touched], deg) +
                        (torch.arange(int(deg.sum()), device=deg.device) -
                         torch.repeat_interleave(torch.cumsum(deg, 0) - deg, deg))] if int(deg.sum()) else \
        torch.zeros(0, dtype=torch.int32, device=deg.device)
    return dict(M=int(sites.numel()), pairs=int(processed.numel()), sum_nadj=int(nadj.sum()),
                sites_touched=int(touched.numel()), neighbours_touched=int(torch.unique(nb).numel()),
                P=width * height)


def algorithmic_bytes(c):
    """Algorithmic HBM bytes per launch. Forward: per listed (tile, site) pair
    the list entry (4 B), the site record (sphere + normal/sigma, 32 B) and its
    8 texels (8 x 28 B), per face 8 B of fp16 encoding + 4 B index, per pixel
    32 B of outputs, per distinct site 8 B of contrib/point_err. Backward: the
    forward's reads plus 36 B of saved/incoming per-pixel data and, per pair,
    a read-modify-write of the site and texel gradients, and of the sphere
    gradients of touched neighbours."""
    read = c["pairs"] * (4 + 32 + 224 + 8) + 12 * c["sum_nadj"]
    fwd = read + c["P"] * 32 + 8 * c["sites_touched"]
    bwd = read + c["P"] * (12 + 4 + 4 + 12 + 4) + c["pairs"] * 2 * (32 + 224) + 2 * 16 * c["neighbours_touched"]
    return dict(raster_fwd=fwd, raster_bwd=bwd)


def cpu_baseline(w, width, height, min_seconds=10.0, max_views=4):
    """FP32 oracle fwd+bwd on the workload's first views (same inputs), one
    frame after another until at least `min_seconds` of CPU work are timed."""
    from oracle import orc
    from powerfoam_amd.synth import sh_eval
    s = w["scene"]
    sp = np.concatenate([s["points"], s["radii"][:, None]], 1).astype(np.float32)
    ns = np.concatenate([s["normals"], s["density"][:, None]], 1).astype(np.float32)
    target = w["target"].cpu().numpy()
    ts = s["texel_sites"].astype(np.float64)
    dt, frames = 0.0, 0
    while frames < min(max_views, len(w["cams"])) and (frames == 0 or dt < min_seconds):
        cam_t = w["cams"][frames]
        cam, _ = orc.camera(cam_t.eye.numpy(), cam_t.right.numpy(), cam_t.up.numpy(), width, height)
        d = ts - cam_t.eye.numpy().astype(np.float64)[None, None]
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        rgb = sh_eval(s["sh"][:, None].astype(np.float64), d).astype(np.float32)
        rgbh = np.concatenate([rgb, s["texel_height"][..., None]], -1)
        t0 = time.time()
        offsets, sites, _ = orc.bin_lists(cam, sp)
        diff = orc.adj_encode(sp, w["adj_np"], w["off_np"])
        inp = orc.Inputs(sp, ns, s["texel_sites"], rgbh, w["adj_np"], w["off_np"], diff)
        o = orc.raster_fwd(cam, inp, offsets, sites, ray_gt=target, return_point_err=True, precision="f32")
        g_color = (2.0 * (o["color"] - target) / o["color"].size).astype(np.float32)
        z = np.zeros((height, width), np.float32)
        orc.raster_bwd(cam, inp, offsets, sites, o["tile_early_stop"], o["log_t"], g_color, z, z,
                       np.zeros((height, width, 3), np.float32), np.zeros(sp.shape[0], np.float32),
                       precision="f32")
        dt += time.time() - t0
        frames += 1
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=frames * width * height / dt / 1e6, unit="Mrays/s", cores=cores, kind="port",
                sample=f"{frames} frame(s) (views 0-{frames - 1}) of the workload: binning + adjacency encode + "
                       f"fwd + bwd, FP32 oracle, {dt:.1f}s",
                host=host_cpu())


def host_cpu():
    """The measuring machine's CPU model and its online core count (nproc)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, nproc {len(os.sched_getaffinity(0))}, os.cpu_count {os.cpu_count()}"


def valu_peak(dev, n_blocks=4096, iters=4096, reps=3):
    """Wave64 VALU instructions per second the GPU issues on independent FP32
    FMA chains (pf_probe_valu), best of `reps` timed launches."""
    import torch
    from powerfoam_amd import _abi
    from powerfoam_amd._abi import ptr, stream_handle
    lib = _abi.load()
    out = torch.empty(n_blocks * 256, dtype=torch.float32, device=dev)
    best = 0.0
    for r in range(reps + 1):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        n = lib.pf_probe_valu(n_blocks, iters, ptr(out), stream_handle())
        e.record()
        if n < 0:
            _abi.check(int(n))
        torch.cuda.synchronize()
        if r > 0:  # the first launch warms up
            best = max(best, n / (s.elapsed_time(e) * 1e-3))
    return best


def timed_views(steps, n_views):
    """View index of each timed frame (frames cycle through the cameras)."""
    return [i % n_views for i in range(steps)]


def per_view_profile(name, kernel, key, views):
    """Mean over the timed views of a per-view figure from a committed
    rocprofv3 counter summary (profiles/<name>), with its source; None if
    the summary is missing."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    vals = d.get("kernels", {}).get(kernel, {}).get(key)
    if not vals:
        return None, None
    return float(np.mean([vals[v % len(vals)] for v in views])), f"profiles/{name}: {d.get('command', '')}"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--sites", type=int, default=1_000_000)
    ap.add_argument("--views", type=int, default=32)
    ap.add_argument("--width", type=int, default=1297)
    ap.add_argument("--height", type=int, default=840)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-trace", action="store_true",
                    help="skip the ray tracer record (1M-site fisheye views) of the default line")
    ap.add_argument("--trace-steps", type=int, default=16)
    ap.add_argument("--no-mlp-head", action="store_true",
                    help="leave the learned-head record (3M-site scene) out of the default line")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run 3 device copies of this many MiB (counter calibration)")
    ap.add_argument("--head", choices=("sh", "mlp"), default="sh",
                    help="mlp: only the learned-head workload (3M-site scene, MLP vs SH head, PSNR)")
    ap.add_argument("--mlp-sites", type=int, default=3_000_000)
    ap.add_argument("--mlp-steps", type=int, default=2000)
    ap.add_argument("--mlp-batch", type=int, default=65536)
    ap.add_argument("--mlp-frames", type=int, default=16)
    ap.add_argument("--mlp-encoding", choices=("freq", "sh"), default="freq",
                    help="direction encoding of the timed MLP head (the other is trained for its PSNR)")
    ap.add_argument("--workload", choices=("raster", "trace", "cech"), default="raster",
                    help="raster: the rasterizer frame (default); trace: the ray tracer on fisheye views; "
                         "cech: the Cech adjacency build")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the renderer hot path is built and timed on a single MI355X; "
              "multi-GPU runs are not measured here; the per-view sharding is data parallel over cameras.",
              file=sys.stderr)
        sys.exit(2)
    if "OMP_NUM_THREADS" not in os.environ:
        os.environ["OMP_NUM_THREADS"] = str(min(16, os.cpu_count() or 1))

    import torch
    from types import SimpleNamespace
    from powerfoam_amd.color_fn import SHColor
    from powerfoam_amd.rasterize import Rasterizer

    dev = torch.device("cuda:0")
    if args.workload == "trace":
        return run_trace(args, dev)
    if args.workload == "cech":
        return run_cech(args, dev)
    if args.head == "mlp":
        rec, ms = mlp_head_record(args, dev)
        P = args.width * args.height
        line = dict(metric=METRIC, value=P / ms / 1e3, unit="Mrays/s", n_gpus=1, steps=args.mlp_frames,
                    warmup=args.warmup, ms_per_step=ms, higher_is_better=True, scaling="weak", vs_baseline=None,
                    dtype="bf16 head / f32 geometry", data="synthetic (seeded scene generator, random target image)",
                    config=dict(workload="rasterizer fwd+bwd with the learned MLP colour head",
                                sites=args.mlp_sites, width=args.width, height=args.height, views=args.views),
                    mlp_head=rec)
        print(json.dumps(line), flush=True)
        return
    w = build_workload(args.sites, args.views, args.width, args.height, dev, seed=1)
    rast = Rasterizer(SimpleNamespace(num_texel_sites=8, render_objective="volume", is_pinhole=True,
                                      disable_coop_prim_load=False, disable_coop_adj_load=False), dev)
    head = SHColor()
    cams = w["cams"]

    if args.calib_copy_mb:
        n = args.calib_copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads and writes n*4 bytes
        torch.cuda.synchronize()
        log(f"calibration: 3 copies of {args.calib_copy_mb} MiB")
        del src, dst

    log(f"warmup {args.warmup} frames")
    for i in range(args.warmup):
        frame(rast, head, w, cams[i % len(cams)])
    torch.cuda.synchronize()

    log(f"timing {args.steps} frames")
    rast.event_log = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        frame(rast, head, w, cams[i % len(cams)])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ev = rast.event_log
    rast.event_log = None
    kern = {}
    for name, s, e in ev:
        kern.setdefault(name, []).append(s.elapsed_time(e))
    kern_ms = {k: float(np.mean(v)) for k, v in kern.items()}
    ms = elapsed * 1e3 / args.steps
    P = args.width * args.height
    value = P * args.steps / elapsed / 1e6
    log(f"{ms:.3f} ms/frame, {value:.1f} Mrays/s; kernels {kern_ms}")

    # workload counts of every view; the roofline averages them over the
    # views the timed frames rendered, as the kernel times do
    per_view = []
    sp_all = torch.cat([w["params"]["points"], w["params"]["radii"][:, None]], 1).detach().float().contiguous()
    for v in range(len(cams)):
        frame(rast, head, w, cams[v])
        per_view.append(workload_counts(rast, w, args.width, args.height))
        # the lists without the culling of sites that cannot shade (what the
        # reference bins)
        per_view[-1]["M_all"] = int(rast.bin(cams[v], sp_all)[0][-1])
    torch.cuda.synchronize()
    views = timed_views(args.steps, len(cams))
    avg = {k: float(np.mean([per_view[v][k] for v in views])) for k in per_view[0]}
    abytes = algorithmic_bytes(avg)
    dom = max(kern_ms, key=kern_ms.get)
    achieved = abytes[dom] / (kern_ms[dom] * 1e-3) / 1e9
    traffic, traffic_src = per_view_profile("raster_c1_traffic.json", dom, "per_view_bytes", views)
    roofline = dict(bound="hbm", kernel=dom, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=traffic, algorithmic_bytes=abytes[dom],
                    traffic_source=traffic_src)
    peak_valu = valu_peak(dev)
    log(f"VALU peak (FMA probe): {peak_valu:.4e} wave64 instructions/s")
    valu = {}
    for k in ("raster_fwd", "raster_bwd"):
        insts, src = per_view_profile("raster_c1_valu.json", k, "per_view_valu_insts", views)
        if insts is None or k not in kern_ms:
            continue
        ach = insts / (kern_ms[k] * 1e-3)
        valu[k] = dict(achieved=ach, peak=peak_valu, unit="wave64 VALU instructions/s", frac=ach / peak_valu,
                       insts_per_launch=insts, source=src)
    roofline["valu"] = valu
    roofline["bound_note"] = ("compositing does ~300-500 VALU wave-instructions per (tile, site) pair against "
                              "~300 B of site data shared by 64 rays: VALU-bound; the HBM fraction is reported "
                              "as asked, the VALU fraction beside it")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (FP32 oracle, frames until 10 s)")
        cpu = cpu_baseline(w, args.width, args.height)
        log(f"cpu baseline: {cpu}")

    line = dict(metric=METRIC, value=value, unit="Mrays/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=ms, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic (seeded scene generator, random target image)",
                config=dict(workload="rasterizer fwd+bwd with SH-3 colour head, 1M-site GMM scene, "
                                     "32 pinhole views 1297x840",
                            sites=args.sites, width=args.width, height=args.height, views=args.views,
                            adjacency="power/alpha", adjacency_edges=int(w["stats"]["edges"]),
                            pairs_per_frame=avg["M"], pairs_per_frame_all_sites=avg["M_all"],
                            processed_pairs_per_frame=avg["pairs"]),
                kernel_ms=kern_ms, roofline=roofline, cpu_baseline=cpu)
    del w, rast
    torch.cuda.empty_cache()
    # the extra records never cost the headline line: an error in one is
    # recorded in its place
    if not args.no_trace:
        try:
            line["trace"] = trace_record(args, dev, args.trace_steps, 2)
        except Exception as e:  # noqa: BLE001
            log(f"ray-tracer record failed: {e!r}")
            line["trace"] = dict(error=repr(e))
        torch.cuda.empty_cache()
    if not args.no_mlp_head:
        try:
            line["mlp_head"], _ = mlp_head_record(args, dev)
        except Exception as e:  # noqa: BLE001
            log(f"learned-head record failed: {e!r}")
            line["mlp_head"] = dict(error=repr(e))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
