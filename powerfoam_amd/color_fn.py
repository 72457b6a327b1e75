"""Texel colour heads on the gfx950 kernels: per camera they turn per-site
parameters into the rasterizer's texel_rgb [N,T,3], with hand-written
backwards and, like the reference's head, no gradient to the texel
positions.

* `SphericalVoronoi` / `SphericalVoronoiGradFn`: the reference's own head
  (powerfoam/color_fn.py:10-315) with its signatures, on pf_sv_fwd /
  pf_sv_bwd.
* `SHColor`: degree-3 SH (pf_sh_fwd / pf_sh_bwd), the head the synthetic
  benchmark scenes carry, in the same slot (called where scene.py:331-335
  calls the SV head).
"""

import ctypes
import math

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


_FOV_BUFFER_RAD = math.radians(10.0)


class SphericalVoronoiGradFn(torch.autograd.Function):
    """color_fn.py:10-24"""

    @staticmethod
    def forward(ctx, spherical_voronoi, points, camera, att_sites, att_values, att_temps):
        fov_cos_cutoff = getattr(camera, "fov_cos_cutoff", None)
        if fov_cos_cutoff is None:
            fov_cos_cutoff = spherical_voronoi.fov_cos_cutoff
        return spherical_voronoi._forward(ctx, points, camera, att_sites, att_values, att_temps, fov_cos_cutoff)

    @staticmethod
    def backward(ctx, *args):
        return ctx.spherical_voronoi._backward(ctx, *args)


class SphericalVoronoi:
    """The reference's Spherical-Voronoi texel colour (color_fn.py:27-315):
    forward(points [P,3], camera, att_sites [D,P,3], att_values [D,P,3],
    att_temps [D,P]) -> texel colours [P,3], differentiable w.r.t. the three
    lobe arrays."""

    @staticmethod
    def compute_fov_cos_cutoff(camera):
        """color_fn.py:28-36 for a pinhole whose principal point is the image
        centre: the frustum diagonal's tangent widened by 10 %. (The forward
        replaces it by the camera's own diagonal plus 10 degrees, as the
        reference does, color_fn.py:218-222.)"""
        right = float(torch.linalg.norm(torch.as_tensor(camera.right)))
        up = float(torch.linalg.norm(torch.as_tensor(camera.up)))
        w, h = float(camera.width), float(camera.height)
        x_max, y_max = right * w / max(w - 1.0, 1.0), up * h / max(h - 1.0, 1.0)
        tan_theta = (x_max ** 2 + y_max ** 2) ** 0.5 * 1.1
        return 1.0 / ((1.0 + tan_theta ** 2) ** 0.5)

    def __init__(self, args, device, attr_dtype="float"):
        self.device = torch.device(device)
        self.args = args
        self.sv_dof = int(args.sv_dof)
        if attr_dtype == "float":
            self.tscalar = torch.float32
        elif attr_dtype == "half":
            self.tscalar = torch.float16
        else:
            raise ValueError(f"Unsupported attribute dtype: {attr_dtype}")
        self.fov_cos_cutoff = None
        self.lib = _abi.load()

    def _args(self, points, camera, att_sites, att_values, att_temps, cutoff):
        a = _abi.PfSvArgs()
        a.n_points = points.shape[0]
        a.sv_dof = att_sites.shape[0]
        a.points = points.data_ptr()
        eye = torch.as_tensor(camera.eye).detach().to("cpu", torch.float64).reshape(3)
        right = torch.as_tensor(camera.right).detach().to("cpu", torch.float64).reshape(3)
        up = torch.as_tensor(camera.up).detach().to("cpu", torch.float64).reshape(3)
        fwd = torch.cross(up, right, dim=-1)
        fwd = fwd / torch.linalg.norm(fwd)
        a.eye[:] = eye.tolist()
        a.forward[:] = fwd.tolist()
        a.fov_cos_cutoff = float(cutoff)
        a.att_sites, a.att_values, a.att_temps = att_sites.data_ptr(), att_values.data_ptr(), att_temps.data_ptr()
        return a

    def _forward(self, grad_ctx, points, camera, att_sites, att_values, att_temps, fov_cos_cutoff):
        d = att_sites.shape[0]
        n = points.shape[0]
        if att_sites.shape != (d, n, 3) or att_values.shape != (d, n, 3) or att_temps.shape != (d, n):
            raise ValueError("att_sites/att_values must be [D,P,3] and att_temps [D,P] for P points")
        # the cut-off from the camera's own frustum diagonal (color_fn.py:218-222)
        right = float(torch.linalg.norm(torch.as_tensor(camera.right).float()))
        up = float(torch.linalg.norm(torch.as_tensor(camera.up).float()))
        fov_cos_cutoff = math.cos(math.atan(math.sqrt(right ** 2 + up ** 2)) + _FOV_BUFFER_RAD)
        ins = [_c32(x, self.tscalar) for x in (att_sites, att_values, att_temps)]
        pts = points.detach().to(torch.float32).contiguous()
        values = torch.empty((n, 3), dtype=torch.float32, device=pts.device)
        wsum = torch.empty(n, dtype=torch.float32, device=pts.device)
        a = self._args(pts, camera, *ins, fov_cos_cutoff)
        check(self.lib.pf_sv_fwd(ctypes.byref(a), ptr(values), ptr(wsum), stream_handle()))
        grad_ctx.spherical_voronoi = self
        grad_ctx.camera = camera
        grad_ctx.fov_cos_cutoff = fov_cos_cutoff
        grad_ctx.dtypes = (att_sites.dtype, att_values.dtype, att_temps.dtype)
        grad_ctx.save_for_backward(pts, *ins, values, wsum)
        return values.to(self.tscalar)

    def forward(self, points, camera, att_sites, att_values, att_temps):
        return SphericalVoronoiGradFn.apply(self, points, camera, att_sites, att_values, att_temps)

    def _backward(self, grad_ctx, grad_values):
        pts, sites, vals, temps, values, wsum = grad_ctx.saved_tensors
        g = grad_values.detach().to(torch.float32).contiguous()
        gs, gv, gt = torch.empty_like(sites), torch.empty_like(vals), torch.empty_like(temps)
        a = self._args(pts, grad_ctx.camera, sites, vals, temps, grad_ctx.fov_cos_cutoff)
        check(self.lib.pf_sv_bwd(ctypes.byref(a), ptr(wsum), ptr(values), ptr(g), ptr(gs), ptr(gv), ptr(gt),
                                 stream_handle()))
        ds, dv, dt = grad_ctx.dtypes
        del grad_ctx.spherical_voronoi, grad_ctx.camera
        return None, None, None, gs.to(ds), gv.to(dv), gt.to(dt)


def _c32(t, tscalar):
    """float32 copy, rounded through the attribute dtype first (fp16 storage)."""
    return t.detach().to(tscalar).to(torch.float32).contiguous()


def get_att_sv(texel_sv_axis, texel_sv_rgb, sv_dof):
    """scene.py:290-312: [N,T,3D] raw lobe parameters -> lobe-major
    (axis [D,NT,3] unit, values [D,NT,3], temps [D,NT] = |raw axis|),
    differentiable (torch ops)."""
    T = texel_sv_axis.shape[1]
    axis = texel_sv_axis.view(-1, T, sv_dof, 3)
    temp = axis.norm(dim=-1)
    axis = axis / axis.norm(dim=-1, keepdim=True)
    axis = axis.permute(2, 0, 1, 3).reshape(sv_dof, -1, 3).contiguous()
    rgb = texel_sv_rgb.view(-1, T, sv_dof, 3).permute(2, 0, 1, 3).reshape(sv_dof, -1, 3).contiguous()
    temp = temp.permute(2, 0, 1).reshape(sv_dof, -1).contiguous()
    return axis, rgb, temp
