"""Learned colour head on BF16 matrix cores (pf_mlp_fwd / pf_mlp_bwd).

No reference counterpart: it fills the texel-colour slot that the reference's
SphericalVoronoi pre-pass fills (color_fn.py:54-184, called from
scene.py:314-337) and that `color_fn.SHColor` fills with degree-3 SH. A row of
site s seen along unit direction d gets

    x = [z_s ; e(d)]                          z_s: per-site 16-dim latent
    rgb = max(0, W3 relu(W2 relu(W1 x + b1) + b2) + b3)

with 32 -> 64 -> 64 -> 3 layers and 16 direction features e(d): the SH head's
basis (1, Y_1(d) .. Y_15(d)), exactly what the SH head contracts with its 16
coefficients per channel, or the frequency encoding (1, d, sin(pi d),
cos(pi d), sin(2 pi d), cos(2 pi d)), per component. The kernels
read bf16 copies of the latents and weights; `MLPColor` keeps FP32 masters
(what the optimiser updates) and makes the bf16 copies again when a master's
version counter moved (an optimiser step, an in-place write).

`MLPColor.forward(texel_sites, camera)` returns texel_rgb [N,T,3] for a
camera and is differentiable w.r.t. the latents and weights (not the texel
positions, as color_fn.py:181-184); its texel-mode backward runs over the
sites whose texels received a colour gradient only. The rasterizer can also
evaluate the head itself (`Rasterizer.forward(..., color_head=head)`): once
per site with a tile in the view, between binning and compositing
(`Rasterizer.head_mode = "binned"`), or inside the compositing forward
(`"compositing"`). `fit` trains the head on the GPU
with the same kernels on (site, direction) pairs labelled by the SH head.
"""

import math

import numpy as np
import torch

from . import _abi
from ._abi import check, stream_handle

LATENT = 16
HIDDEN = 64
PARAM_NAMES = ("latents", "w1", "b1", "w2", "b2", "w3", "b3")


def _ref(struct):
    import ctypes
    return ctypes.byref(struct)


class MLPFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, head, rows_key, rows_args, *params):
        ctx.head, ctx.rows_args = head, rows_args
        return head._run_fwd(rows_args)

    @staticmethod
    def backward(ctx, grad_rgb):
        grads = ctx.head._run_bwd(ctx.rows_args, grad_rgb)
        del ctx.head, ctx.rows_args
        return (None, None, None) + grads


ENCODINGS = {"sh": _abi.MLP_ENC_SH, "freq": _abi.MLP_ENC_FREQ}


class MLPColor:
    def __init__(self, n_sites, device, seed=0, latents=None, encoding="sh"):
        """Weights drawn with He initialisation from `seed`; latents given or
        zero. encoding: the 16 direction features, "sh" (1, Y_1..Y_15, the SH
        head's basis) or "freq" (1, d, sin(pi d), cos(pi d), sin(2 pi d),
        cos(2 pi d) per component)."""
        if encoding not in ENCODINGS:
            raise ValueError(f"encoding must be one of {sorted(ENCODINGS)}")
        self.encoding = encoding
        self.lib = _abi.load()
        self.device = torch.device(device)
        g = torch.Generator(device="cpu").manual_seed(seed)
        f32 = dict(dtype=torch.float32)

        def he(o, i):
            return (torch.randn((o, i), generator=g, **f32) * math.sqrt(2.0 / i)).to(self.device)

        self.params = dict(
            latents=(torch.zeros((n_sites, LATENT), **f32) if latents is None
                     else torch.as_tensor(latents, **f32).clone()).to(self.device).contiguous(),
            w1=he(HIDDEN, 2 * LATENT), b1=torch.zeros(HIDDEN, device=self.device),
            w2=he(HIDDEN, HIDDEN), b2=torch.zeros(HIDDEN, device=self.device),
            w3=he(3, HIDDEN) * 0.5, b3=torch.zeros(3, device=self.device))
        for p in self.params.values():
            p.requires_grad_(True)
        self.event_log = None
        self._bf16_cache, self._bf16_stamp = None, None
        # texel-mode backward over the sites with a colour gradient only
        # (pf_mlp_active_sites); False: every row
        self.compact_backward = True

    # ------------------------------------------------------------- plumbing
    def _bf16(self):
        """bf16 copies the kernels read (16-byte aligned, contiguous), made
        again only when an FP32 master changed since (its version counter,
        which in-place optimiser steps and assignments advance)."""
        keys = ("latents", "w1", "w2", "w3")
        stamp = tuple((self.params[k].data_ptr(), self.params[k]._version) for k in keys)
        if self._bf16_stamp != stamp:
            self._bf16_cache = {k: self.params[k].detach().to(torch.bfloat16).contiguous() for k in keys}
            self._bf16_stamp = stamp
        return self._bf16_cache

    def refresh(self):
        """Drop the bf16 copies: for writes that bypass the version counter
        (through `.data`); in-place writes and optimiser steps need no call."""
        self._bf16_cache, self._bf16_stamp = None, None

    def _weights(self, bf):
        w = _abi.PfMlpWeights()
        w.w1, w.w2, w.w3 = bf["w1"].data_ptr(), bf["w2"].data_ptr(), bf["w3"].data_ptr()
        for k in ("b1", "b2", "b3"):
            setattr(w, k, self.params[k].detach().contiguous().data_ptr())
        return w

    def _rows(self, a, bf):
        r = _abi.PfMlpRows()
        r.n_rows = a["n_rows"]
        r.n_sites = self.params["latents"].shape[0]
        r.n_texels = a.get("T", 1)
        r.texel_sites = a["texel_sites"].data_ptr() if "texel_sites" in a else None
        r.eye[:] = a.get("eye", (0.0, 0.0, 0.0))
        r.pair_sites = a["pair_sites"].data_ptr() if "pair_sites" in a else None
        r.pair_dirs = a["pair_dirs"].data_ptr() if "pair_dirs" in a else None
        r.latents = bf["latents"].data_ptr()
        r.encoding = ENCODINGS[self.encoding]
        return r

    def _launch(self, name, fn, *a):
        if self.event_log is None:
            return check(fn(*a))
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        st = fn(*a)
        e.record()
        self.event_log.append((name, s, e))
        return check(st)

    def _run_fwd(self, a):
        bf = self._bf16()
        out = torch.empty((a["n_rows"], 3), dtype=torch.float32, device=self.device)
        w, r = self._weights(bf), self._rows(a, bf)
        self._launch("mlp_fwd", self.lib.pf_mlp_fwd, _ref(w), _ref(r), out.data_ptr(), stream_handle())
        return out

    def _run_bwd(self, a, grad_rgb, site_list=None):
        """site_list: (list, count) device tensors naming every site whose
        rows can carry a colour gradient (texel mode), or None."""
        bf = self._bf16()
        g = {k: torch.zeros_like(self.params[k]) for k in PARAM_NAMES}
        gs = _abi.PfMlpGrads(*[g[k].data_ptr() for k in PARAM_NAMES])
        w, r = self._weights(bf), self._rows(a, bf)
        gr = grad_rgb.detach().to(torch.float32).contiguous()
        n_sites = self.params["latents"].shape[0]
        if site_list is not None:
            r.site_list, r.site_count = site_list[0].data_ptr(), site_list[1].data_ptr()
        elif "texel_sites" in a and self.compact_backward and a["n_rows"] == n_sites * a.get("T", 1):
            # only the sites whose texels got a colour gradient (the rows of
            # the others add nothing): a device list, no host sync
            lst = torch.empty(max(n_sites, 1), dtype=torch.int32, device=self.device)
            cnt = torch.zeros(1, dtype=torch.int32, device=self.device)
            check(self.lib.pf_mlp_active_sites(gr.data_ptr(), n_sites, a.get("T", 1), lst.data_ptr(), cnt.data_ptr(),
                                               stream_handle()))
            r.site_list, r.site_count = lst.data_ptr(), cnt.data_ptr()
        self._launch("mlp_bwd", self.lib.pf_mlp_bwd, _ref(w), _ref(r), gr.data_ptr(), _ref(gs), stream_handle())
        return tuple(g[k] for k in PARAM_NAMES)

    def _apply(self, a):
        return MLPFn.apply(self, None, a, *[self.params[k] for k in PARAM_NAMES])

    # ------------------------------------------------------------------ API
    def forward(self, texel_sites, camera):
        """texel_rgb [N,T,3] for `camera` (its eye), from texel_sites [N,T,3]."""
        n, T = texel_sites.shape[0], texel_sites.shape[1]
        ts = texel_sites.detach().to(torch.float32).contiguous()
        eye = torch.as_tensor(camera.eye).detach().to("cpu", torch.float32).reshape(-1)
        a = dict(n_rows=n * T, T=T, texel_sites=ts, eye=tuple(float(v) for v in eye))
        return self._apply(a).view(n, T, 3)

    def forward_pairs(self, sites, dirs):
        """rgb [B,3] of (site, direction) pairs."""
        a = dict(n_rows=int(sites.shape[0]), pair_sites=sites.to(torch.int32).contiguous(),
                 pair_dirs=dirs.detach().to(torch.float32).contiguous())
        return self._apply(a)

    def parameters(self):
        return [self.params[k] for k in PARAM_NAMES]


def sh_pair_colors(sh, sites, dirs):
    """SH-head colour of (site, direction) pairs through pf_sh_fwd: each pair is
    a one-texel site at unit distance from an eye at the origin."""
    lib = _abi.load()
    b = int(sites.shape[0])
    shp = sh[sites.long()].to(torch.float32).contiguous()
    ts = dirs.to(torch.float32).reshape(b, 1, 3).contiguous()
    out = torch.empty((b, 1, 3), dtype=torch.float32, device=sh.device)
    import ctypes
    eye = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    check(lib.pf_sh_fwd(shp.data_ptr(), ts.data_ptr(), b, 1, eye, out.data_ptr(), stream_handle()))
    return out.view(b, 3)


def latents_from_sh(sh, seed=0):
    """Initial latents: the DC colour (3) and a fixed random projection of the
    45 higher-band coefficients (13), scaled to the DC's range."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    p = torch.randn((45, LATENT - 3), generator=g) / math.sqrt(45.0)
    hi = sh[:, 1:, :].reshape(sh.shape[0], 45).to(torch.float32)
    return torch.cat([sh[:, 0, :].to(torch.float32), (hi @ p.to(sh.device)) * 10.0], dim=1).contiguous()


def fit(head, sh, steps=2000, batch=65536, lr=3e-3, lr_latent=3e-2, seed=0, log_every=0):
    """Train `head` on (site, uniform direction) pairs labelled by the SH head,
    L2 loss, Adam, with the project's MLP kernels for forward and backward.
    Returns the list of (step, loss)."""
    dev = head.device
    n = sh.shape[0]
    g = torch.Generator(device=dev).manual_seed(seed)
    opt = torch.optim.Adam([{"params": [head.params["latents"]], "lr": lr_latent},
                            {"params": [head.params[k] for k in PARAM_NAMES[1:]], "lr": lr}])
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    hist = []
    for it in range(steps):
        sites = torch.randint(0, n, (batch,), device=dev, generator=g, dtype=torch.int64).to(torch.int32)
        d = torch.randn((batch, 3), device=dev, generator=g)
        d = d / d.norm(dim=1, keepdim=True)
        with torch.no_grad():
            target = sh_pair_colors(sh, sites, d)
        pred = head.forward_pairs(sites, d)
        loss = ((pred - target) ** 2).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if log_every and (it % log_every == 0 or it == steps - 1):
            hist.append((it, float(loss.detach())))
    return hist


def psnr(a, b, peak=1.0):
    mse = float(((a - b) ** 2).mean())
    return float("inf") if mse == 0 else 10.0 * math.log10(peak * peak / mse)
