"""Cameras: the reference's TorchCamera (powerfoam/camera.py:181-215) without Warp.

`TorchCamera` keeps the reference's fields and `get_ray_dir` convention
(camera.py:19-24): for pixel (i, j) the ray is x*right + y*up + forward with
x = 2j/(W-1) - 1, y = 1 - 2i/(H-1) and forward = normalize(up x right).
`to_pf()` replaces `to_warp()`: it builds the C-ABI camera struct, and unlike
the reference it never rebuilds pinhole ray maps (pinhole rays are generated
inside the kernels).
"""

import math
from dataclasses import dataclass

import numpy as np
import torch

from ._abi import PfCamera


@dataclass
class TorchCamera:
    eye: torch.Tensor
    right: torch.Tensor
    up: torch.Tensor
    width: int
    height: int
    ray_maps: torch.Tensor = None
    fov_cos_cutoff: float = None

    def _vec(self, v):
        v = torch.as_tensor(v).detach().to("cpu", torch.float32).reshape(-1)
        return [float(v[0]), float(v[1]), float(v[2])]

    def to_pf(self):
        cam = PfCamera()
        cam.eye[:] = self._vec(self.eye)
        cam.right[:] = self._vec(self.right)
        cam.up[:] = self._vec(self.up)
        cam.width = int(self.width)
        cam.height = int(self.height)
        if self.ray_maps is not None:
            rm = self.ray_maps
            if rm.dtype != torch.float32 or not rm.is_contiguous():
                raise ValueError("ray_maps must be a contiguous float32 [H,W,6] tensor")
            if tuple(rm.shape) != (self.height, self.width, 6):
                raise ValueError(f"ray_maps shape {tuple(rm.shape)} != {(self.height, self.width, 6)}")
            cam.ray_maps = rm.data_ptr()
        else:
            cam.ray_maps = None
        cam.fov_cos_cutoff = float(self.fov_cos_cutoff) if self.fov_cos_cutoff is not None else 0.0
        return cam

    @property
    def is_pinhole(self):
        return self.ray_maps is None

    def get_ray_dir(self, i, j):
        """camera.py:19-24, vectorised over pixel indices (unnormalised)."""
        right = torch.as_tensor(self.right, dtype=torch.float32, device=i.device)
        up = torch.as_tensor(self.up, dtype=torch.float32, device=i.device)
        fwd = torch.linalg.cross(up, right)
        fwd = fwd / torch.linalg.norm(fwd)
        x = 2.0 * j / (float(self.width) - 1.0) - 1.0
        y = 1.0 - 2.0 * i / (float(self.height) - 1.0)
        return x[:, None] * right[None] + y[:, None] * up[None] + fwd[None]

    def _build_pinhole_ray_maps(self):
        dev = torch.as_tensor(self.eye).device
        i = torch.arange(self.height, dtype=torch.float32, device=dev)
        j = torch.arange(self.width, dtype=torch.float32, device=dev)
        ii, jj = torch.meshgrid(i, j, indexing="ij")
        d = self.get_ray_dir(ii.reshape(-1), jj.reshape(-1))
        d = d / torch.linalg.norm(d, dim=-1, keepdim=True)
        o = torch.as_tensor(self.eye, dtype=torch.float32, device=dev)[None].expand(d.shape[0], 3)
        return torch.cat([o, d], dim=-1).view(self.height, self.width, 6)


def look_at(eye, target, width, height, fov_deg, world_up=(0.0, 1.0, 0.0), device="cpu"):
    """Pinhole camera with a horizontal field of view of `fov_deg` measured to
    the edge pixel centres: |right| = tan(fov/2), |up| = |right| (H-1)/(W-1),
    and up x right = forward (the reference's convention)."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    r = np.cross(fwd, np.asarray(world_up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, fwd)
    tr = math.tan(math.radians(fov_deg) / 2.0)
    right = r * tr
    up = u * tr * (height - 1) / (width - 1)
    t = lambda v: torch.tensor(np.asarray(v, np.float32), device=device)
    return TorchCamera(t(eye), t(right), t(up), int(width), int(height))


def fisheye_ray_maps(eye, target, size, fov_deg=180.0, world_up=(0.0, 1.0, 0.0)):
    """Equidistant fisheye: a pixel at radius rho from the image centre looks at
    angle theta = (fov/2) * rho / rho_max off the optical axis, rho_max being half
    the image width (corners continue past fov/2). Returns float32 [H,W,6]."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    r = np.cross(fwd, np.asarray(world_up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, fwd)
    c = (size - 1) / 2.0
    ii, jj = np.meshgrid(np.arange(size, dtype=np.float64), np.arange(size, dtype=np.float64), indexing="ij")
    dx, dy = jj - c, c - ii
    rho = np.sqrt(dx * dx + dy * dy)
    theta = math.radians(fov_deg) / 2.0 * rho / (size / 2.0)
    phi = np.arctan2(dy, dx)
    s = np.sin(theta)
    d = (np.cos(theta)[..., None] * fwd + (s * np.cos(phi))[..., None] * r + (s * np.sin(phi))[..., None] * u)
    o = np.broadcast_to(eye, d.shape)
    return np.concatenate([o, d], axis=-1).astype(np.float32)
