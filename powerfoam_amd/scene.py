"""Scene glue on the gfx950 kernels (pf_glue_fwd / pf_glue_bwd): the map from
a Power Foam scene's trainable parameters to the rasterizer's per-site
inputs, as the reference's PowerfoamScene computes it before every render
(powerfoam/scene.py:239-337), fused into one launch with a hand-written
adjoint.

`SceneGlue` holds the parameters under the reference's names (points, radii,
quaternions, density, texel_sites -- the per-texel (tangent, bitangent)
offsets in units of the radius --, texel_height) and offers
`get_density()`, `get_radii()`, `get_normals()` and `rasterizer_inputs()`;
the colour head supplies texel_rgb separately.
"""

import torch

from . import _abi
from ._abi import check, stream_handle

_NAMES = ("points", "radii", "density", "quaternions", "texel_offsets", "texel_height")
_OUT = ("radii", "density", "normals", "texel_sites", "texel_height")


def _c(t):
    return t.detach().to(torch.float32).contiguous()


class GlueFn(torch.autograd.Function):
    """(points, radii_raw, density_raw, quaternions, texel_offsets, height_raw)
    -> (radii, density, normals, texel_sites, texel_height)."""

    @staticmethod
    def forward(ctx, points, radii_raw, density_raw, quaternions, texel_offsets, height_raw):
        lib = _abi.load()
        n, T = texel_offsets.shape[0], texel_offsets.shape[1]
        if (points.shape != (n, 3) or radii_raw.shape != (n,) or density_raw.shape != (n,)
                or quaternions.shape != (n, 4) or texel_offsets.shape != (n, T, 2) or height_raw.shape != (n, T)):
            raise ValueError("points [N,3], radii/density [N], quaternions [N,4], texel offsets [N,T,2] and "
                             "texel heights [N,T] must agree")
        ins = [_c(x) for x in (points, radii_raw, density_raw, quaternions, texel_offsets, height_raw)]
        dev = points.device
        f32 = dict(dtype=torch.float32, device=dev)
        outs = dict(radii=torch.empty(n, **f32), density=torch.empty(n, **f32), normals=torch.empty((n, 3), **f32),
                    texel_sites=torch.empty((n, T, 3), **f32), texel_height=torch.empty((n, T), **f32))
        p = _abi.PfGlueParams(*[x.data_ptr() for x in ins])
        o = _abi.PfGlueOutputs(*[outs[k].data_ptr() for k in _OUT])
        check(lib.pf_glue_fwd(_ref(p), n, T, _ref(o), stream_handle()))
        ctx.save_for_backward(*ins)
        ctx.set_materialize_grads(False)
        return tuple(outs[k] for k in _OUT)

    @staticmethod
    def backward(ctx, *grads):
        lib = _abi.load()
        ins = ctx.saved_tensors
        n, T = ins[4].shape[0], ins[4].shape[1]
        gs = [None if g is None else _c(g) for g in grads]
        gp = {k: torch.empty_like(x) for k, x in zip(_NAMES, ins)}
        p = _abi.PfGlueParams(*[x.data_ptr() for x in ins])
        g = _abi.PfGlueOutputs(*[None if x is None else x.data_ptr() for x in gs])
        out = _abi.PfGlueGrads(*[gp[k].data_ptr() for k in _NAMES])
        check(lib.pf_glue_bwd(_ref(p), n, T, _ref(g), _ref(out), stream_handle()))
        return tuple(gp[k] for k in _NAMES)


def _ref(struct):
    import ctypes
    return ctypes.byref(struct)


class SceneGlue:
    """The trainable parameters of a Power Foam scene (scene.py:100-180 names)
    and the rasterizer inputs derived from them (scene.py:239-337)."""

    def __init__(self, points, radii, quaternions, density, texel_sites, texel_height, device="cuda",
                 requires_grad=True):
        f = lambda t: torch.as_tensor(t, dtype=torch.float32).detach().clone().to(device).contiguous() \
            .requires_grad_(requires_grad)
        self.points = f(points)
        self.radii = f(radii)
        self.quaternions = f(quaternions)
        self.density = f(density)
        self.texel_sites = f(texel_sites)
        self.texel_height = f(texel_height)

    def parameters(self):
        return [self.points, self.radii, self.quaternions, self.density, self.texel_sites, self.texel_height]

    def _glue(self):
        return GlueFn.apply(self.points, self.radii, self.density, self.quaternions, self.texel_sites,
                            self.texel_height)

    def get_density(self):
        return self._glue()[1]

    def get_radii(self):
        return self._glue()[0]

    def get_normals(self):
        return self._glue()[2]

    def rasterizer_inputs(self):
        """dict(points, radii, density, normals, texel_sites, texel_height) for
        Rasterizer.forward, differentiable w.r.t. the parameters."""
        radii, density, normals, texel_sites, texel_height = self._glue()
        return dict(points=self.points, radii=radii, density=density, normals=normals, texel_sites=texel_sites,
                    texel_height=texel_height)
