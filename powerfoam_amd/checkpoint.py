"""Reader of the reference's trained scenes (`model.pt`, written by
PowerfoamScene.save_pt, powerfoam/scene.py:707-731) into the drop-in
renderer's inputs.

The file is a torch dict of CPU tensors: points [N,3], density [N], radii [N]
(both before softplus), quaternions [N,4], texel_sites [N,T,2] (tangent-plane
offsets in units of the radius), texel_sv_axis and texel_sv_rgb [N,T,3 D]
(the Spherical-Voronoi lobes), texel_height [N,T] (in units of the radius),
adjacency [E] and adjacency_offsets [N+1]. It is read with
torch.load(weights_only=True), which executes nothing from the file; the
reference's load_pt (scene.py:733-756) casts every attribute but the points
to the scene's attribute dtype, as `attr_dtype` does here.

`PowerfoamCheckpoint.render_inputs(camera, sv_head)` gives the arguments of
Rasterizer.forward after texel_rgb, the way PowerfoamScene.forward builds
them (scene.py:314-353): the scene glue on the GPU (pf_glue_fwd) and the
Spherical-Voronoi colour of each texel site for the camera (pf_sv_fwd),
differentiable w.r.t. every stored parameter.
"""

import torch

from .color_fn import get_att_sv
from .scene import SceneGlue

KEYS = ("points", "density", "radii", "quaternions", "texel_sites", "texel_sv_axis", "texel_sv_rgb", "texel_height",
        "adjacency", "adjacency_offsets")


class PowerfoamCheckpoint:
    def __init__(self, data, device="cuda", attr_dtype="float", requires_grad=False):
        missing = [k for k in KEYS if k not in data]
        if missing:
            raise ValueError(f"model.pt lacks {missing}")
        n = data["points"].shape[0]
        T = data["texel_sites"].shape[1]
        if data["texel_sv_axis"].shape[-1] % 3 or data["texel_sv_axis"].shape != data["texel_sv_rgb"].shape:
            raise ValueError("texel_sv_axis and texel_sv_rgb must both be [N,T,3*sv_dof]")
        if data["adjacency_offsets"].shape != (n + 1,):
            raise ValueError("adjacency_offsets must hold N+1 offsets")
        self.n_sites, self.n_texels = n, T
        self.sv_dof = data["texel_sv_axis"].shape[-1] // 3
        self.device = torch.device(device)
        ts = {"float": torch.float32, "half": torch.float16}[attr_dtype]
        q = lambda t: t.to(ts).to(torch.float32)
        self.glue = SceneGlue(data["points"].float(), q(data["radii"]), q(data["quaternions"]), q(data["density"]),
                              q(data["texel_sites"]), q(data["texel_height"]), device=self.device,
                              requires_grad=requires_grad)
        f = lambda t: q(t).to(self.device).contiguous().requires_grad_(requires_grad)
        self.texel_sv_axis = f(data["texel_sv_axis"])
        self.texel_sv_rgb = f(data["texel_sv_rgb"])
        self.adjacency = data["adjacency"].to(torch.int32).to(self.device).contiguous()
        self.adjacency_offsets = data["adjacency_offsets"].to(torch.int32).to(self.device).contiguous()

    @classmethod
    def load(cls, path, device="cuda", attr_dtype="float", requires_grad=False):
        data = torch.load(path, map_location="cpu", weights_only=True)
        if not isinstance(data, dict):
            raise ValueError(f"{path}: not a scene dict")
        return cls(data, device, attr_dtype, requires_grad)

    def parameters(self):
        return self.glue.parameters() + [self.texel_sv_axis, self.texel_sv_rgb]

    def render_inputs(self, camera, sv_head):
        """dict(points, radii, density, normals, texel_sites, texel_rgb,
        texel_height, adjacency, adjacency_offsets) for Rasterizer.forward."""
        x = self.glue.rasterizer_inputs()
        axis, rgb, temp = get_att_sv(self.texel_sv_axis, self.texel_sv_rgb, self.sv_dof)
        texel_rgb = sv_head.forward(x["texel_sites"].reshape(-1, 3).detach(), camera, axis, rgb, temp)
        x["texel_rgb"] = texel_rgb.view(self.n_sites, self.n_texels, 3).to(torch.float32)
        x["adjacency"] = self.adjacency
        x["adjacency_offsets"] = self.adjacency_offsets
        return x


def save_pt(path, points, density, radii, quaternions, texel_sites, texel_sv_axis, texel_sv_rgb, texel_height,
            adjacency, adjacency_offsets):
    """Write a scene in the reference's model.pt layout (scene.py:707-731)."""
    c = lambda t: torch.as_tensor(t).detach().float().cpu()
    torch.save(dict(points=c(points), density=c(density), radii=c(radii), quaternions=c(quaternions),
                    texel_sites=c(texel_sites), texel_sv_axis=c(texel_sv_axis), texel_sv_rgb=c(texel_sv_rgb),
                    texel_height=c(texel_height), adjacency=torch.as_tensor(adjacency).cpu(),
                    adjacency_offsets=torch.as_tensor(adjacency_offsets).cpu()), path)
