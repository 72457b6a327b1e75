"""Drop-in replacement for the reference's powerfoam/raytrace.py (RayTracer) on
gfx950 kernels (pf_trace_fwd / pf_trace_bwd).

`RayTracer.benchmark` keeps the reference signature and output
(raytrace.py:247-309). Additions, behind new names only: `forward` returns
colour, opacity, depth and normals and is differentiable (RayTraceGradFn, the
exact adjoint -- the reference has no backward), `trace_cells` returns the
visited cells per ray, and `start_cell` picks the start cell the way
benchmark.py:320-324 does. Non-pinhole cameras (ray_maps) are supported.
"""

import torch

from . import _abi
from ._abi import check, stream_handle
from .rasterize import _f32c, _ref, _scene_struct

TILE_WIDTH = 8
TILE_SIZE = 64


def start_cell(points, radii, eye):
    """argmin_i |eye - c_i|^2 - r_i^2: the power cell containing the eye (benchmark.py:320-324)."""
    eye = torch.as_tensor(eye, dtype=points.dtype, device=points.device)
    dists = torch.linalg.norm(points - eye[None, :], dim=-1)
    return int(torch.argmin(dists ** 2 - radii ** 2))


class RayTraceGradFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tracer, *args):
        return tracer._forward(ctx, *args)

    @staticmethod
    def backward(ctx, *grads):
        return ctx.tracer._backward(ctx, *grads)


class RayTracer:
    def __init__(self, args, device, attr_dtype="float"):
        if attr_dtype == "float":
            self.tscalar = torch.float32
        elif attr_dtype == "half":
            self.tscalar = torch.float16
        else:
            raise ValueError(f"Unsupported attribute dtype: {attr_dtype}")
        self.device = torch.device(device)
        self.args = args
        self.attr_dtype = attr_dtype
        self.num_texel_sites = int(args.num_texel_sites)
        if not 1 <= self.num_texel_sites <= _abi.MAX_TEXELS:
            raise ValueError(f"num_texel_sites must be in 1..{_abi.MAX_TEXELS}")
        self.lib = _abi.load()
        self.max_steps = 1 << 16
        self.event_log = None
        # Ballot compaction of the forward walk over groups of this many
        # neighbouring 8x8 tiles (0: one wave per tile). Results are identical.
        # Off by default: on the 1M-site fisheye views a tile's lanes already
        # do 90 % useful steps (its 64 rays end within a few cells of each
        # other), while grouping tiles divides the waves in flight of a
        # latency-bound walk: forward 2.26 ms per view without, 3.92 ms with
        # groups of 2 tiles, 6.36 ms with groups of 4. It also leaves the
        # backward without the per-tile walk lengths of its longest-first order.
        self.compact = 0
        # The backward takes its tiles longest walk first (pf_trace_tile_order
        # over the per-tile walk lengths the forward records), so the waves of
        # the longest walks start early instead of running alone at the end.
        # Needs compact = 0. The gradients' summation order follows the order.
        self.longest_first = True
        # With exact geometry (no adjacency_diff) the walk reads each face's
        # plane from one record written per call (pf_trace_planes: the same
        # FP32 values it would form from the neighbour's sphere) instead of
        # the neighbour's id and sphere. Results are identical either way.
        self.face_planes = True

    def _launch(self, name, fn, *a):
        if self.event_log is None:
            return check(fn(*a))
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        st = fn(*a)
        e.record()
        self.event_log.append((name, s, e))
        return check(st)

    def _prep(self, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
              adjacency_offsets, adjacency_diff):
        q = lambda t: t.detach().to(self.tscalar).to(torch.float32).contiguous()
        all_spheres = _f32c(torch.cat([points, radii[:, None]], dim=-1))
        all_nsigmas = q(torch.cat([normals, density[:, None]], dim=-1))
        all_texel_sites = q(texel_sites)
        all_texel_rgbh = torch.cat([q(texel_rgb), q(texel_height)[..., None]], dim=-1).contiguous()
        adjacency = adjacency.to(torch.int32).contiguous()
        adjacency_offsets = adjacency_offsets.to(torch.int32).contiguous()
        diff = adjacency_diff.to(torch.float16).contiguous() if adjacency_diff is not None else None
        if texel_sites.shape[1] != self.num_texel_sites:
            raise ValueError("texel_sites does not match args.num_texel_sites")
        planes = None
        if diff is None and self.face_planes and adjacency.numel() > 0:
            planes = torch.empty((adjacency.numel(), 4), dtype=torch.float32, device=all_spheres.device)
            sc = _scene_struct(all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency, adjacency_offsets,
                               None)
            check(self.lib.pf_trace_planes(_ref(sc), adjacency.numel(), planes.data_ptr(), stream_handle()))
        return all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency, adjacency_offsets, diff, planes

    def _args(self, camera, start_point_idx, tensors, threshold, depth_quantile):
        all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency, adjacency_offsets, diff, planes = tensors
        a = _abi.PfTraceArgs()
        a.cam = camera.to_pf()
        a.scene = _scene_struct(all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency,
                                adjacency_offsets, diff)
        if isinstance(start_point_idx, torch.Tensor) and start_point_idx.numel() > 1:
            ss = start_point_idx.to(torch.int32).contiguous()
            a.start_sites = ss.data_ptr()
            a.start_site = 0
            keep = ss
        else:
            a.start_sites = None
            a.start_site = int(start_point_idx)
            keep = None
        a.transmittance_threshold = float(threshold)
        a.depth_quantile = float(depth_quantile)
        a.max_steps = int(self.max_steps)
        a.compact = int(self.compact)
        a.face_planes = planes.data_ptr() if planes is not None else None
        return a, keep

    def benchmark(self, camera, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                  texel_height, adjacency, adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3):
        """raytrace.py:247-309: colour of every pixel's cell walk."""
        with torch.no_grad():
            t = self._prep(points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                           adjacency_offsets, adjacency_diff)
            a, keep = self._args(camera, start_point_idx, t, transmittance_threshold, 0.5)
            color = torch.zeros((camera.height, camera.width, 3), dtype=torch.float32, device=points.device)
            a.color = color.data_ptr()
            self._launch("trace_fwd", self.lib.pf_trace_fwd, _ref(a), stream_handle())
            return color

    def trace_cells(self, camera, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                    texel_height, adjacency, adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3,
                    max_cells=64):
        """Visited cells per ray ([H,W,max_cells], -1 padded) and their count [H,W]."""
        with torch.no_grad():
            t = self._prep(points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                           adjacency_offsets, adjacency_diff)
            a, keep = self._args(camera, start_point_idx, t, transmittance_threshold, 0.5)
            H, W = camera.height, camera.width
            color = torch.zeros((H, W, 3), dtype=torch.float32, device=points.device)
            cells = torch.full((H, W, max_cells), -1, dtype=torch.int32, device=points.device)
            n_cells = torch.zeros((H, W), dtype=torch.int32, device=points.device)
            a.color, a.cells, a.max_cells, a.n_cells = color.data_ptr(), cells.data_ptr(), max_cells, n_cells.data_ptr()
            check(self.lib.pf_trace_fwd(_ref(a), stream_handle()))
            return cells, n_cells

    def forward(self, camera, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                texel_height, adjacency, adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3,
                depth_quantile=0.5):
        """(color [H,W,3], opacity [H,W], depth [H,W] (-1 where T never drops below
        depth_quantile), normal [H,W,3]); differentiable w.r.t. points, radii, density,
        normals, texel_sites, texel_rgb, texel_height. adjacency_diff=None uses exact
        neighbour geometry."""
        return RayTraceGradFn.apply(self, camera, start_point_idx, points, radii, density, normals, texel_sites,
                                    texel_rgb, texel_height, adjacency, adjacency_offsets, adjacency_diff,
                                    transmittance_threshold, depth_quantile)

    def _forward(self, ctx, camera, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                 texel_height, adjacency, adjacency_offsets, adjacency_diff, transmittance_threshold,
                 depth_quantile):
        t = self._prep(points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                       adjacency_offsets, adjacency_diff)
        a, keep = self._args(camera, start_point_idx, t, transmittance_threshold, depth_quantile)
        H, W = camera.height, camera.width
        f32 = dict(dtype=torch.float32, device=points.device)
        color = torch.zeros((H, W, 3), **f32)
        log_t = torch.zeros((H, W), **f32)
        normal = torch.zeros((H, W, 3), **f32)
        depth = torch.zeros((H, W), **f32)
        q_step = torch.zeros((H, W), dtype=torch.int32, device=points.device)
        q_sigma = torch.zeros((H, W), **f32)
        a.color, a.log_t, a.normal, a.depth = color.data_ptr(), log_t.data_ptr(), normal.data_ptr(), depth.data_ptr()
        a.q_step, a.q_sigma = q_step.data_ptr(), q_sigma.data_ptr()
        tile_steps = None
        if self.longest_first and self.compact == 0:
            n_tiles = -(-H // TILE_WIDTH) * -(-W // TILE_WIDTH)
            tile_steps = torch.empty(n_tiles, dtype=torch.int32, device=points.device)
            a.tile_steps = tile_steps.data_ptr()
        self._launch("trace_fwd", self.lib.pf_trace_fwd, _ref(a), stream_handle())
        ctx.tile_steps = tile_steps
        ctx.tracer = self
        ctx.camera = camera
        ctx.start = start_point_idx if keep is None else None
        ctx.params = (transmittance_threshold, depth_quantile)
        ctx.save_for_backward(*t[:6], *(x for x in t[6:8] if x is not None), color, normal, log_t, q_step, q_sigma,
                              *(x for x in (keep,) if x is not None))
        ctx.has_diff = t[6] is not None
        ctx.has_planes = t[7] is not None
        ctx.has_starts = keep is not None
        opacity = 1.0 - torch.exp(log_t)
        return color, opacity, depth, normal

    def _backward(self, ctx, g_color, g_opacity, g_depth, g_normal):
        saved = list(ctx.saved_tensors)
        k = 6
        diff = saved[k] if ctx.has_diff else None
        k += ctx.has_diff
        planes = saved[k] if ctx.has_planes else None
        k += ctx.has_planes
        tensors = saved[:6] + [diff, planes]
        rest = saved[k:]
        color, normal, log_t, q_step, q_sigma = rest[:5]
        starts = rest[5] if ctx.has_starts else ctx.start
        camera = ctx.camera
        H, W = camera.height, camera.width
        dev = color.device
        z = lambda *s: torch.zeros(*s, dtype=torch.float32, device=dev)
        g = lambda x, shape: _f32c(x) if x is not None else z(shape)
        gC, gO, gD, gN = g(g_color, (H, W, 3)), g(g_opacity, (H, W)), g(g_depth, (H, W)), g(g_normal, (H, W, 3))
        a, keep = self._args(camera, starts, tensors, *ctx.params)
        N, T = tensors[0].shape[0], tensors[2].shape[1]
        gs, gn, gts, gtr, gth = z((N, 4)), z((N, 4)), z((N, T, 3)), z((N, T, 3)), z((N, T))
        b = _abi.PfTraceBwdArgs()
        b.fwd = a
        b.color, b.normal, b.log_t = color.data_ptr(), normal.data_ptr(), log_t.data_ptr()
        b.q_step, b.q_sigma = q_step.data_ptr(), q_sigma.data_ptr()
        b.grad_color, b.grad_normal, b.grad_depth, b.grad_opacity = gC.data_ptr(), gN.data_ptr(), gD.data_ptr(), \
            gO.data_ptr()
        b.grad_spheres, b.grad_nsigmas = gs.data_ptr(), gn.data_ptr()
        b.grad_texel_sites, b.grad_texel_rgb, b.grad_texel_height = gts.data_ptr(), gtr.data_ptr(), gth.data_ptr()
        st = stream_handle()
        order = None
        if ctx.tile_steps is not None:
            order = torch.empty_like(ctx.tile_steps)
            check(self.lib.pf_trace_tile_order(ctx.tile_steps.data_ptr(), ctx.tile_steps.numel(), order.data_ptr(), st))
            b.tile_order = order.data_ptr()
        self._launch("trace_bwd", self.lib.pf_trace_bwd, _ref(b), st)
        for x in (gs, gn, gts, gtr, gth):
            check(self.lib.pf_scrub_nonfinite(_abi.ptr(x), x.numel(), st))
        del ctx.tracer, ctx.camera, ctx.tile_steps
        return (None, None, None, gs[:, :3], gs[:, 3], gn[:, 3], gn[:, :3], gts, gtr, gth, None, None, None, None,
                None)
