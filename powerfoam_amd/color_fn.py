"""Degree-3 SH colour head on the gfx950 kernels (pf_sh_fwd / pf_sh_bwd).

Stands where the reference's SphericalVoronoi colour pre-pass stands
(powerfoam/color_fn.py:27-315, called from scene.py:331-335): per camera it
turns per-site parameters into the rasterizer's texel_rgb [N,T,3], with a
hand-written backward and, like the reference head, no gradient to the texel
positions.
"""

import ctypes

import torch

from . import _abi
from ._abi import check, ptr, stream_handle


class SHGradFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sh, texel_sites, eye):
        lib = _abi.load()
        sh_c = sh.detach().to(torch.float32).contiguous()
        ts = texel_sites.detach().to(torch.float32).contiguous()
        n, T = ts.shape[0], ts.shape[1]
        if sh_c.shape != (n, 16, 3):
            raise ValueError(f"sh must be [N,16,3], got {tuple(sh_c.shape)}")
        rgb = torch.empty((n, T, 3), dtype=torch.float32, device=ts.device)
        e = (ctypes.c_float * 3)(*[float(v) for v in eye])
        check(lib.pf_sh_fwd(ptr(sh_c), ptr(ts), n, T, e, ptr(rgb), stream_handle()))
        ctx.save_for_backward(sh_c, ts)
        ctx.eye = tuple(float(v) for v in eye)
        return rgb

    @staticmethod
    def backward(ctx, grad_rgb):
        lib = _abi.load()
        sh_c, ts = ctx.saved_tensors
        g = grad_rgb.detach().to(torch.float32).contiguous()
        grad_sh = torch.empty_like(sh_c)
        e = (ctypes.c_float * 3)(*ctx.eye)
        check(lib.pf_sh_bwd(ptr(sh_c), ptr(ts), ptr(g), ts.shape[0], ts.shape[1], e, ptr(grad_sh),
                            stream_handle()))
        return grad_sh, None, None


class SHColor:
    """texel_rgb = SHColor().forward(texel_sites, camera, sh)"""

    def forward(self, texel_sites, camera, sh):
        eye = torch.as_tensor(camera.eye).detach().cpu().reshape(-1).tolist()
        return SHGradFn.apply(sh, texel_sites, eye)
