"""Drop-in replacement for the reference's powerfoam/rasterize.py (Rasterizer,
RasterGradFn, VisOptions) on hand-written gfx950 kernels.

Same class names, constructor and method signatures, outputs, gradient tuple
and error behaviour as the reference (rasterize.py:15-65, 1587-1980); all
launches go through the C-ABI library (powerfoam_amd/_abi.py) on torch's
current stream. There is no CPU path: without the HIP library every call
raises.
"""

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from ._abi import check, ptr, stream_handle
from .camera import TorchCamera

TILE_WIDTH = 8
TILE_SIZE = TILE_WIDTH * TILE_WIDTH
LEAF_SIZE = 2


@dataclass
class VisOptions:
    """rasterize.py:15-20. Fields start at zero like the reference's wp.struct;
    Rasterizer.visualize(vis_options=None) uses threshold 1e-3, 1024
    intersections, a black background and depth quantile 0 (:2146-2150)."""
    transmittance_threshold: float = 0.0
    max_intersections: int = 0
    depth_quantile: float = 0.0
    bkgd_color: tuple = field(default_factory=lambda: (0.0, 0.0, 0.0))


class RasterGradFn(torch.autograd.Function):
    """rasterize.py:23-34"""

    @staticmethod
    def forward(ctx, rasterizer, *args):
        ctx.n_args = len(args)
        # outputs the loss does not use come back as None (no zero tensors
        # materialised per output); the backward reads one shared zero block
        ctx.set_materialize_grads(False)
        return rasterizer._forward(ctx, *args)

    @staticmethod
    def backward(ctx, *args):
        n_args = ctx.n_args
        grads = ctx.rasterizer._backward(ctx, *args)
        # one entry per forward input (the rasterizer itself included); an
        # explicit transmittance_threshold gets None like the other non-tensors
        return tuple(grads) + (None,) * (1 + n_args - len(grads))


def _f32c(t):
    return t.detach().to(torch.float32).contiguous()


class Rasterizer:
    def __init__(self, args, device, attr_dtype="float"):
        """rasterize.py:38-65. `args` needs render_objective, num_texel_sites and
        is_pinhole (the reference's Params fields); disable_coop_* are accepted
        and ignored (loads are always staged through LDS)."""
        if attr_dtype == "float":
            self.tscalar = torch.float32
        elif attr_dtype == "half":
            self.tscalar = torch.float16
        else:
            raise ValueError(f"Unsupported attribute dtype: {attr_dtype}")
        self.device = torch.device(device)
        self.args = args
        self.attr_dtype = attr_dtype
        self.loss_accumulation = getattr(args, "render_objective", "volume") == "surface"
        self.num_texel_sites = int(args.num_texel_sites)
        if not 1 <= self.num_texel_sites <= _abi.MAX_TEXELS:
            raise ValueError(f"num_texel_sites must be in 1..{_abi.MAX_TEXELS}")
        self.lib = _abi.load()
        # Not in the reference: derivative-check switches. exact_geometry feeds
        # the faces full-precision neighbour spheres instead of the fp16
        # encoding; exact_gradients differentiates the forward exactly (see
        # pf_raster_bwd_args.exact_gradients).
        self.exact_geometry = False
        self.exact_gradients = False
        # Optional timing: when a list, (name, start, end) torch.cuda.Events are
        # appended around each kernel launch, recorded on the launch stream.
        self.event_log = None
        # Tile lists and early-stop counters of the last forward, for workload
        # accounting (the same tensors the backward keeps).
        self.last_tiles = None
        # Start the tiles with the longest site lists first (the results do not
        # depend on the order; the kernels' tail gets shorter).
        self.longest_first = True
        # Split site lists longer than this many sites into segments that run
        # in parallel (None: one wave per list). A multiple of 64. The forward
        # is bound by its longest lists and gains when they are long (see
        # split_min); the backward is bound by its total work, so by default
        # it walks whole lists (split_backward).
        self.seg_len = 512
        self.split_backward = False
        # Split only when some list is longer than this many sites. Splitting
        # adds a log-T pass over every segment but the last of each split
        # list, a launch of its own before the main pass. On the 32 measured
        # views of the 1M-site scene (longest lists 1,034-2,314 sites once the
        # sites that cannot shade are left out; long lists mostly reach the
        # transmittance stop before their end) whole lists are faster on every
        # view; a lone wave walks about 0.6 sites per microsecond, so lists of
        # several thousand walked sites would set the forward's duration.
        self.split_min = 4096
        self._longest = None
        self._dropped = None
        # Leave sites that no ray from the camera can shade (their sphere lies
        # beyond one of their own faces: pf_live_sites) out of the tile lists
        # of the forward, backward and benchmark passes. Results are the same
        # up to float summation order; the visualisation pass, whose
        # intersection counts follow list positions, always bins every site.
        self.cull_hidden = True
        # ... and bin each remaining site only to tiles that meet a ball
        # holding everything it can shade (its sphere cut down by its faces).
        self.bound_sites = True
        # The forward marks the list entries that shade some ray of their tile
        # and the backward walks only those (an entry that shades nothing adds
        # nothing to any gradient). Not in the reference.
        self.skip_unshaded = True
        # True: the compositing tests only the faces that can bound a ray's
        # interval inside their site's ball (pf_prune_faces: a face whose disc
        # in the ball lies outside another face, or whose plane misses the
        # ball, is left out; margin * radius of slack covers float rounding).
        # The cells inside the balls are unchanged. Off by default: on the
        # 1M-site views the compositing kernels ran no faster with the pruned
        # faces (their time is not in the face loop) and the pruning pass
        # costs ~0.6 ms per frame. Not in the reference.
        self.prune_faces = False
        self.prune_margin = 1e-3
        # "bf16": the compositing kernels read the texel colour and height
        # features as bfloat16 (8 B per texel instead of 16); geometry,
        # transmittance and every gradient stay FP32. The backward treats the
        # rounding as the identity (the gradient of the stored feature is the
        # gradient of the FP32 input). Not in the reference.
        self.feature_dtype = "float"
        # With a color_head: keep the packed texel rows of the last forward
        # (rgb written by the fused head for the sites that shade, height) in
        # head_rows, for inspection. Off: the rows live only until backward.
        self.keep_head_rows = False
        self.head_rows = None
        # Where a color_head runs: "binned" (default) evaluates it once per
        # site with a tile in this view (pf_mask_sites over the binning's
        # visibility mask, then a texel-mode pf_mlp_fwd writing the colours
        # into the packed rows) between binning and compositing;
        # "compositing" evaluates it inside the compositing forward for each
        # (tile, site) that shades (pf_raster_fwd_args.mlp_head). Same rows,
        # same outputs.
        self.head_mode = "binned"

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

    # ------------------------------------------------------------------ binning
    def _camera(self, camera, ray_maps=None):
        """(C-ABI camera, tensor to keep alive). With args.is_pinhole the rays
        are generated in the kernels and camera.ray_maps is not read, as in the
        reference's pinhole kernels (rasterize.py:940-944); otherwise they come
        from camera.ray_maps, or from pinhole ray maps built as TorchCamera.to_warp
        does (camera.py:197-215), and binning goes through the cone tree."""
        if getattr(self.args, "is_pinhole", True):
            return TorchCamera(camera.eye, camera.right, camera.up, camera.width, camera.height).to_pf(), None
        rm = ray_maps if ray_maps is not None else camera.ray_maps
        if rm is None:
            rm = TorchCamera(camera.eye, camera.right, camera.up, camera.width, camera.height) \
                ._build_pinhole_ray_maps()
        rm = rm.detach().to(device=self.device, dtype=torch.float32).contiguous()
        return TorchCamera(camera.eye, camera.right, camera.up, camera.width, camera.height, rm).to_pf(), rm

    def bin(self, camera, all_spheres, live=None, bounds=None, checks=()):
        """count -> scan -> (one host sync for M) -> emit -> per-tile sort; ray-map
        cameras count and emit through the cone tree (rasterize.py:1518-1760).
        `live` (uint8 per site, from live_sites) leaves the sites marked 0 out
        of the lists, `bounds` (their balls) the tiles that miss a site's
        ball; the visibility mask still covers both. `checks`: (device scalar,
        expected int, message) triples read in the same host sync; a mismatch
        raises ValueError. Records that an emit pass could not place (its
        offsets disagreeing with the count, rasterize.py:358) are counted on
        the device and raise PfError at the next sync (check_binning()
        forces one)."""
        lib = self.lib
        cam, keep = self._camera(camera)
        th, tw = _abi.tiles_hw(camera.height, camera.width)
        nt = th * tw
        n = all_spheres.shape[0]
        dev = all_spheres.device
        st = stream_handle()
        cones = None
        if keep is not None:
            nodes = lib.pf_cone_tree_nodes(camera.width, camera.height)
            if nodes < 0:
                raise ValueError("ray-map cameras need 5..8192 pixels a side")
            cones = torch.empty((nodes, 4), dtype=torch.float32, device=dev)
            check(lib.pf_bin_cones(_ref(cam), ptr(cones), st))
        counts = torch.empty(nt + 1, dtype=torch.int32, device=dev)
        if cones is None:
            check(lib.pf_bin_count(_ref(cam), ptr(all_spheres), ptr(live), ptr(bounds), n, ptr(counts), st))
        else:
            check(lib.pf_bin_count_cones(_ref(cam), ptr(cones), ptr(all_spheres), n, ptr(counts), st))
        offsets = torch.empty(nt + 1, dtype=torch.int64, device=dev)
        check(lib.pf_bin_scan(ptr(counts), nt, ptr(offsets), st))
        # the one host sync of the frame: the record count, the longest list
        # (which decides whether lists are split into segments), the records
        # earlier emits dropped, and the caller's consistency checks
        dropped = self._dropped_word(dev)
        vals = torch.stack([offsets[-1], counts.max().to(torch.int64), dropped[0].to(torch.int64)] +
                           [t.reshape(()).to(torch.int64) for t, _, _ in checks]).tolist()
        m, self._longest = int(vals[0]), int(vals[1])
        self._raise_dropped(int(vals[2]))
        for (_, want, msg), got in zip(checks, vals[3:]):
            if int(got) != int(want):
                raise ValueError(f"{msg} (got {int(got)}, expected {int(want)})")
        records = torch.empty(max(m, 1), dtype=torch.int64, device=dev)
        cursor = torch.empty(nt, dtype=torch.int32, device=dev)
        visible = torch.zeros(n, dtype=torch.uint8, device=dev)
        if cones is None:
            check(lib.pf_bin_emit(_ref(cam), ptr(all_spheres), ptr(live), ptr(bounds), n, ptr(offsets), nt, ptr(records),
                                  ptr(cursor), ptr(visible), ptr(dropped), st))
        else:
            check(lib.pf_bin_emit_cones(_ref(cam), ptr(cones), ptr(all_spheres), n, ptr(offsets), nt, ptr(records),
                                        ptr(cursor), ptr(visible), ptr(dropped), st))
        sites = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
        ws_bytes = int(lib.pf_bin_sort_workspace_size(m, nt))
        ws = torch.empty(max(ws_bytes // 4, 1), dtype=torch.int32, device=dev)
        check(lib.pf_bin_sort(ptr(offsets), nt, ptr(records), ptr(sites), ptr(ws), ws_bytes, st))
        return offsets, sites[:m] if m > 0 else sites[:0], visible.bool(), nt

    def _dropped_word(self, dev):
        """Device int32 counting records the emit passes could not place."""
        if self._dropped is None or self._dropped.device != dev:
            self._dropped = torch.zeros(1, dtype=torch.int32, device=dev)
        return self._dropped

    def _raise_dropped(self, n):
        if n:
            self._dropped.zero_()
            raise _abi.PfError(f"binning dropped {n} (tile, site) records: an emit pass found its tile's segment "
                               "full (count and emit disagreed); that frame's lists were incomplete")

    def check_binning(self):
        """Synchronises and raises PfError if an emit pass since the last
        check dropped records (the reference drops them silently, rasterize.py:358)."""
        if self._dropped is not None:
            self._raise_dropped(int(self._dropped.item()))

    def tile_order(self, offsets):
        """Tiles by descending list length (int32, pf_tile_schedule), or None for
        the natural order."""
        if not self.longest_first:
            return None
        nt = offsets.shape[0] - 1
        order = torch.empty(nt, dtype=torch.int32, device=offsets.device)
        check(self.lib.pf_tile_schedule(ptr(offsets), nt, 0, 0, 0, ptr(order), None, None, None, None,
                                        stream_handle()))
        return order

    def segments(self, offsets, n_records, longest=None):
        """Segment table of the compositing kernels (pf_segments) for lists of
        n_records sites in all, with the longest-first tile order under
        "tile_order", built on the device in one launch; or None (splitting
        off, or no list longer than split_min)."""
        if not self.seg_len or (longest is not None and longest <= self.split_min):
            return None
        L = int(self.seg_len)
        nt = offsets.shape[0] - 1
        dev = offsets.device
        cap = nt + int(n_records) // L
        i32 = dict(dtype=torch.int32, device=dev)
        sg = dict(n_segs=cap, seg_len=L, count=torch.empty(1, **i32), tile_first=torch.empty(nt + 1, **i32),
                  seg_tile=torch.empty(cap, **i32), order=torch.empty(cap, **i32),
                  tile_order=torch.empty(nt, **i32) if self.longest_first else None,
                  log_t_sums=torch.empty(cap * 64, dtype=torch.float32, device=dev),
                  lane_active=torch.empty(cap * 64, dtype=torch.uint8, device=dev))
        to = sg["tile_order"]
        check(self.lib.pf_tile_schedule(ptr(offsets), nt, int(n_records), L, cap, ptr(to) if to is not None else None,
                                        ptr(sg["tile_first"]), ptr(sg["seg_tile"]), ptr(sg["order"]),
                                        ptr(sg["count"]), stream_handle()))
        return sg

    def _segs(self, sg, scratch=None):
        out = _abi.PfSegments()
        if sg is None:
            return out
        out.n_segs, out.seg_len = sg["n_segs"], sg["seg_len"]
        out.count = sg["count"].data_ptr()
        out.tile_first, out.seg_tile = sg["tile_first"].data_ptr(), sg["seg_tile"].data_ptr()
        out.order, out.log_t_sums = sg["order"].data_ptr(), sg["log_t_sums"].data_ptr()
        out.lane_active = sg["lane_active"].data_ptr()
        out.scratch = scratch.data_ptr() if scratch is not None else None
        return out

    def _tiles(self, nt, offsets, tile_sites, order):
        return _abi.PfTileLists(nt, offsets.data_ptr(), tile_sites.data_ptr() if tile_sites.numel() else None,
                                order.data_ptr() if order is not None else None)

    def live_sites(self, camera, all_spheres, adjacency, adjacency_offsets, diff, kind=_abi.PF_FACES_RADIUS):
        """(live, bounds): uint8 per site, 0 where the site cannot shade any ray
        from the camera (its sphere lies beyond one of its own faces, or a
        bound ball's cap vanishes), and with `bound_sites` a [N,4] ball per
        site holding everything it can shade (pf_live_sites); faces as the
        compositing kernels read them (`diff` in the encoding `kind`, or the
        full-precision spheres when diff is None). (None, None) when culling
        is off, for ray-map cameras, or without faces."""
        if not self.cull_hidden or adjacency.shape[0] == 0:
            return None, None
        cam, ray_maps = self._camera(camera)
        if ray_maps is not None:
            return None, None
        n = all_spheres.shape[0]
        live = torch.empty(n, dtype=torch.uint8, device=all_spheres.device)
        bounds = torch.empty((n, 4), dtype=torch.float32, device=all_spheres.device) if self.bound_sites else None
        check(self.lib.pf_live_sites(_ref(cam), ptr(all_spheres), n, ptr(adjacency), ptr(adjacency_offsets),
                                     int(adjacency.shape[0]), ptr(diff) if diff is not None else None, int(kind),
                                     ptr(live), ptr(bounds), stream_handle()))
        return live, bounds

    def _features(self, all_texel_rgbh):
        """The bfloat16 copy of the texel features the kernels read (or None)."""
        if self.feature_dtype == "float":
            return None
        if self.feature_dtype != "bf16":
            raise ValueError(f"feature_dtype must be 'float' or 'bf16', not {self.feature_dtype!r}")
        return all_texel_rgbh.to(torch.bfloat16).contiguous()

    def _head_binned(self, head, camera, all_texel_sites, all_texel_rgbh, visible, live):
        """The learned head's texel colours for the sites the tile lists can
        hold (a tile in this view and, with culling, live), written into their
        packed (rgb, height) rows: pf_mask_sites over the visibility and live
        masks, then pf_mlp_fwd over that site list (device-side count, no host
        sync)."""
        n, T = all_texel_rgbh.shape[0], all_texel_rgbh.shape[1]
        dev = all_texel_rgbh.device
        lst = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)
        st = stream_handle()
        check(self.lib.pf_mask_sites(ptr(visible.view(torch.uint8)), ptr(live) if live is not None else None, n,
                                     ptr(lst), ptr(cnt), st))
        bf = head._bf16()
        eye = torch.as_tensor(camera.eye).detach().to("cpu", torch.float32).reshape(-1)
        r = head._rows(dict(n_rows=n * T, T=T, texel_sites=all_texel_sites, eye=tuple(float(v) for v in eye)), bf)
        r.site_list, r.site_count, r.out_stride = ptr(lst), ptr(cnt), 4
        w = head._weights(bf)
        head._launch("mlp_fwd", self.lib.pf_mlp_fwd, _ref(w), _ref(r), ptr(all_texel_rgbh), st)
        return lst, cnt

    def _head_shading_sites(self, head_sites, contrib):
        """The list for the head's backward, in the buffers of the forward's
        list: the sites with contrib > 0. A texel row's colour gradient is a
        sum over rays of alpha*T times the pixel gradient, and contrib[site]
        sums alpha*T/(H*W) over the same rays, so a site off this list has a
        zero row gradient -- unless every alpha*T it met was below 1.5e-39*H*W
        (FP32 underflow of contrib), a gradient under 2e-33 times the pixel
        gradient that the backward then leaves out. The sites binned but hidden
        behind others (about half of those on the forward's list) drop out."""
        lst, cnt = head_sites
        cnt.zero_()
        shades = (contrib != 0).view(torch.uint8)
        check(self.lib.pf_mask_sites(ptr(shades), None, contrib.shape[0], ptr(lst), ptr(cnt), stream_handle()))
        return lst, cnt

    def pack_sites(self, points, radii, normals, density, texel_rgb, texel_height):
        """(spheres [N,4], nsigmas [N,4], texel_rgbh [N,T,4]) in FP32, the
        concatenations of rasterize.py:1619-1624, in one launch (pf_pack_sites).
        texel_rgb None: colours 0 (a learned head fused into the compositing
        supplies them)."""
        n = points.shape[0]
        t = texel_height.shape[1] if texel_height.dim() == 2 else -1
        if (points.shape != (n, 3) or radii.shape != (n,) or normals.shape != (n, 3) or density.shape != (n,)
                or (texel_rgb is not None and texel_rgb.shape != (n, t, 3)) or texel_height.shape != (n, t)):
            raise ValueError("points/normals [N,3], radii/density [N], texel_rgb [N,T,3] and texel_height [N,T] "
                             "must agree")
        dev = points.device
        spheres = torch.empty((n, 4), dtype=torch.float32, device=dev)
        nsigmas = torch.empty((n, 4), dtype=torch.float32, device=dev)
        rgbh = torch.empty((n, t, 4), dtype=torch.float32, device=dev)
        ins = [_f32c(x) if x is not None else None for x in (points, radii, normals, density, texel_rgb, texel_height)]
        check(self.lib.pf_pack_sites(*[ptr(x) for x in ins], n, t, ptr(spheres), ptr(nsigmas), ptr(rgbh),
                                     stream_handle()))
        return spheres, nsigmas, rgbh

    def prune(self, all_spheres, adjacency, adjacency_offsets, adjacency_diff):
        """pf_prune_faces: (adjacency, fp16 faces or None, faces per site) with
        each site's faces that can bound a ray inside its ball first."""
        n, m = all_spheres.shape[0], adjacency.shape[0]
        dev = all_spheres.device
        adj_out = torch.empty(max(m, 1), dtype=torch.int32, device=dev)
        diff_out = torch.empty((max(m, 1), 4), dtype=torch.float16, device=dev) if adjacency_diff is not None else None
        count = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        check(self.lib.pf_prune_faces(ptr(all_spheres), n, ptr(adjacency) if m else None, ptr(adjacency_offsets), m,
                                      ptr(adjacency_diff) if (adjacency_diff is not None and m) else None,
                                      float(self.prune_margin), ptr(adj_out),
                                      ptr(diff_out) if diff_out is not None else None, ptr(count), stream_handle()))
        return adj_out[:m], (diff_out[:m] if diff_out is not None else None), count[:n]

    def encode_adjacency(self, all_spheres, adjacency, adjacency_offsets):
        """prefetch_adjacency_kernel (rasterize.py:580-606): fp16 (c_adj - c, r_adj)."""
        diff = torch.empty((max(adjacency.shape[0], 1), 4), dtype=torch.float16, device=all_spheres.device)
        check(self.lib.pf_adj_encode(ptr(all_spheres), all_spheres.shape[0], ptr(adjacency),
                                     ptr(adjacency_offsets), int(adjacency.shape[0]), ptr(diff), stream_handle()))
        return diff[: adjacency.shape[0]]

    # ------------------------------------------------------------------ forward
    def _forward(self, grad_ctx, camera, depth_quantiles, points, radii, density, normals, texel_sites,
                 texel_rgb, texel_height, adjacency, adjacency_offsets, ray_gt, return_point_err,
                 transmittance_threshold=1e-3, color_head=None, *head_params):
        if self.attr_dtype != "float":
            raise ValueError("Forward pass only supports float precision")
        if ray_gt is None and self.loss_accumulation:
            raise ValueError("Ground truth rays must be provided for surface loss accumulation")
        if ray_gt is None and return_point_err:
            raise ValueError("Ground truth rays must be provided to return point error")
        if texel_sites.shape[1] != self.num_texel_sites:
            raise ValueError("texel_sites does not match args.num_texel_sites")
        if color_head is None and texel_rgb is None:
            raise ValueError("texel_rgb is required without a color_head")
        if color_head is not None and color_head.params["latents"].shape[0] != points.shape[0]:
            raise ValueError("color_head has a latent per site: its latents must match points")
        dev = points.device
        H, W = camera.height, camera.width
        N = points.shape[0]
        all_spheres, all_nsigmas, all_texel_rgbh = self.pack_sites(points, radii, normals, density, texel_rgb,
                                                                   texel_height)
        all_texel_sites = _f32c(texel_sites)
        adjacency = adjacency.to(torch.int32).contiguous()
        adjacency_offsets = adjacency_offsets.to(torch.int32).contiguous()
        ray_gt_c = _f32c(ray_gt) if ray_gt is not None else None
        dq = _f32c(depth_quantiles) if depth_quantiles is not None else None

        adjacency_diff = self.encode_adjacency(all_spheres, adjacency, adjacency_offsets)
        rgbh_bf16 = self._features(all_texel_rgbh)
        live, bounds = self.live_sites(camera, all_spheres, adjacency, adjacency_offsets,
                                       None if self.exact_geometry else adjacency_diff)
        if adjacency_offsets.shape != (N + 1,):
            raise ValueError("adjacency_offsets must hold N+1 offsets")
        offsets, tile_sites, prim_visible_mask, nt = self.bin(
            camera, all_spheres, live, bounds,
            checks=[(adjacency_offsets[-1], adjacency.shape[0], "adjacency_offsets[-1] must equal len(adjacency)")])
        sg = self.segments(offsets, tile_sites.shape[0], self._longest)
        order = self.tile_order(offsets) if sg is None else None
        if color_head is not None and self.head_mode not in ("binned", "compositing"):
            raise ValueError(f"head_mode must be 'binned' or 'compositing', not {self.head_mode!r}")
        head_binned = color_head is not None and self.head_mode == "binned"
        head_sites = None
        if head_binned:
            head_sites = self._head_binned(color_head, camera, all_texel_sites, all_texel_rgbh, prim_visible_mask,
                                           live)
            if rgbh_bf16 is not None:
                # bf16 storage: the heights as stored, the head's colours in FP32
                all_texel_rgbh[..., 3] = rgbh_bf16[..., 3].to(torch.float32)
                rgbh_bf16 = None
        # the compositing's faces: each site's faces that can bound a ray first
        face_count = None
        if self.prune_faces:
            adjacency, pdiff, face_count = self.prune(all_spheres, adjacency, adjacency_offsets,
                                                      None if self.exact_geometry else adjacency_diff)
            if pdiff is not None:
                adjacency_diff = pdiff

        f32 = dict(dtype=torch.float32, device=dev)
        color_out = torch.zeros((H, W, 3), **f32)
        log_t_out = torch.zeros((H, W), **f32)
        normal_distance_out = torch.zeros((H, W), **f32)
        normal_out = torch.zeros((H, W, 3), **f32)
        contrib_out = torch.zeros(N, **f32)
        quantile_depths_out = -torch.ones(dq.shape, **f32) if dq is not None else None
        point_err_out = torch.zeros(N, **f32) if return_point_err else None
        err_out = torch.zeros((H, W), **f32) if self.loss_accumulation else None
        tile_early_stop = torch.zeros(nt, dtype=torch.int32, device=dev)
        entry_shades = (torch.zeros(max(tile_sites.shape[0], 1), dtype=torch.uint8, device=dev)
                        if self.skip_unshaded else None)

        a = _abi.PfRasterFwdArgs()
        a.cam, ray_maps = self._camera(camera)
        a.scene = _scene_struct(all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency,
                                adjacency_offsets, None if self.exact_geometry else adjacency_diff, rgbh_bf16,
                                face_count)
        a.tiles = self._tiles(nt, offsets, tile_sites, order)
        scratch = torch.empty(sg["n_segs"] * 64, **f32) if sg is not None else None
        a.segs = self._segs(sg, scratch)
        a.depth_quantiles = dq.data_ptr() if dq is not None else None
        a.n_quantiles = dq.shape[-1] if dq is not None else 0
        a.ray_gt = ray_gt_c.data_ptr() if ray_gt_c is not None else None
        a.transmittance_threshold = float(transmittance_threshold)
        a.return_point_err = int(bool(return_point_err))
        a.loss_accumulation = int(self.loss_accumulation)
        a.tile_early_stop = tile_early_stop.data_ptr()
        a.color = color_out.data_ptr()
        a.log_t = log_t_out.data_ptr()
        a.normal_distance = normal_distance_out.data_ptr()
        a.normal = normal_out.data_ptr()
        a.quantile_depths = quantile_depths_out.data_ptr() if quantile_depths_out is not None else None
        a.err = err_out.data_ptr() if err_out is not None else None
        a.contrib = contrib_out.data_ptr()
        a.point_err = point_err_out.data_ptr() if point_err_out is not None else None
        a.entry_shades = entry_shades.data_ptr() if entry_shades is not None else None
        head_bf16 = None
        if color_head is not None and not head_binned:
            # the head's colours are computed in the compositing and written,
            # with the heights, into the packed rows the backward reads
            head_bf16 = color_head._bf16()
            hs = _abi.PfMlpHead()
            hs.weights = color_head._weights(head_bf16)
            hs.latents = head_bf16["latents"].data_ptr()
            hs.encoding = _head_encoding(color_head)
            hs.texel_rgbh_out = all_texel_rgbh.data_ptr()
            a.mlp_head = ctypes.addressof(hs)
        self._launch("raster_fwd", self.lib.pf_raster_fwd, _ref(a), stream_handle())
        del head_bf16
        if head_binned:
            head_sites = self._head_shading_sites(head_sites, contrib_out)
        self.head_rows = all_texel_rgbh if (color_head is not None and self.keep_head_rows) else None
        self.last_tiles = (offsets, tile_sites, tile_early_stop)

        grad_ctx.rasterizer = self
        grad_ctx.camera = camera
        grad_ctx.ray_maps = ray_maps
        grad_ctx.order = sg["tile_order"] if sg is not None else order
        grad_ctx.segs = sg
        grad_ctx.entry_shades = entry_shades
        # with the head, the backward reads the colours the forward wrote (FP32)
        grad_ctx.rgbh_bf16 = rgbh_bf16 if color_head is None else None
        grad_ctx.color_head = color_head
        grad_ctx.head_sites = head_sites
        grad_ctx.face_count = face_count
        grad_ctx.save_for_backward(dq, all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency,
                                   adjacency_offsets, adjacency_diff, tile_sites, offsets, tile_early_stop,
                                   ray_gt_c, color_out, log_t_out)
        opacity_out = 1.0 - torch.exp(log_t_out)
        return (color_out, opacity_out, normal_distance_out, normal_out, quantile_depths_out, err_out,
                contrib_out, point_err_out, prim_visible_mask)

    def forward(self, *args, color_head=None):
        """rasterize.py:1858-1859. Not in the reference: color_head, a
        `mlp_head.MLPColor` evaluated on the matrix cores inside the
        compositing forward for the sites that shade (texel_rgb is then not
        read and may be None); its latents and weights get gradients through
        the head's backward over the texel colour gradients."""
        if color_head is None:
            return RasterGradFn.apply(self, *args)
        if len(args) == 13:
            args = args + (1e-3,)
        return RasterGradFn.apply(self, *args, color_head, *color_head.parameters())

    # ---------------------------------------------------------------- benchmark
    def benchmark(self, camera, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                  adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3):
        """rasterize.py:1982-2131: colour only, per-ray stop at T < threshold,
        faces from the (c_adj - c, pm_adj - pm) fp16 encoding of benchmark.py
        (see power_face_encoding). Attributes are rounded through fp16 when the
        rasterizer was built with attr_dtype="half", as the reference stores them."""
        with torch.no_grad():
            H, W = camera.height, camera.width
            q = (lambda t: t.to(self.tscalar).to(torch.float32).contiguous())
            all_spheres = _f32c(torch.cat([points, radii[:, None]], dim=-1))
            all_nsigmas = q(torch.cat([normals, density[:, None]], dim=-1))
            all_texel_sites = q(texel_sites)
            all_texel_rgbh = torch.cat([q(texel_rgb), q(texel_height)[..., None]], dim=-1).contiguous()
            adjacency = adjacency.to(torch.int32).contiguous()
            adjacency_offsets = adjacency_offsets.to(torch.int32).contiguous()
            adjacency_diff = adjacency_diff.to(torch.float16).contiguous()
            live, bounds = self.live_sites(camera, all_spheres, adjacency, adjacency_offsets, adjacency_diff,
                                           _abi.PF_FACES_POWER)
            offsets, tile_sites, _, nt = self.bin(camera, all_spheres, live, bounds)
            order = self.tile_order(offsets)
            color_out = torch.zeros((H, W, 3), dtype=torch.float32, device=points.device)
            a = _abi.PfRasterRgbArgs()
            a.cam, ray_maps = self._camera(camera)
            a.scene = _scene_struct(all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency,
                                    adjacency_offsets, adjacency_diff)
            a.tiles = self._tiles(nt, offsets, tile_sites, order)
            a.transmittance_threshold = float(transmittance_threshold)
            a.color = color_out.data_ptr()
            self._launch("raster_rgb", self.lib.pf_raster_rgb, _ref(a), stream_handle())
            return color_out

    # ------------------------------------------------------------- visualize
    def visualize(self, camera, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                  adjacency_offsets, vis_options=None):
        """rasterize.py:2133-2311: (color [H,W,3] over vis_options.bkgd_color,
        depth [H,W] where T first drops below depth_quantile (0 if never),
        normal [H,W,3], alpha [H,W], intersections [H,W] int32). Attributes go
        through fp16 when attr_dtype="half", as the reference stores them."""
        if vis_options is None:
            vis_options = VisOptions(transmittance_threshold=1e-3, max_intersections=1024)
        with torch.no_grad():
            H, W = camera.height, camera.width
            dev = points.device
            q = (lambda t: t.to(self.tscalar).to(torch.float32).contiguous())
            all_spheres = _f32c(torch.cat([points, radii[:, None]], dim=-1))
            all_nsigmas = q(torch.cat([normals, density[:, None]], dim=-1))
            all_texel_sites = q(texel_sites)
            all_texel_rgbh = torch.cat([q(texel_rgb), q(texel_height)[..., None]], dim=-1).contiguous()
            adjacency = adjacency.to(torch.int32).contiguous()
            adjacency_offsets = adjacency_offsets.to(torch.int32).contiguous()
            offsets, tile_sites, _, nt = self.bin(camera, all_spheres)
            order = self.tile_order(offsets)
            adjacency_diff = self.encode_adjacency(all_spheres, adjacency, adjacency_offsets)
            f32 = dict(dtype=torch.float32, device=dev)
            color = torch.zeros((H, W, 3), **f32)
            depth = torch.zeros((H, W), **f32)
            normal = torch.zeros((H, W, 3), **f32)
            alpha = torch.zeros((H, W), **f32)
            inter = torch.zeros((H, W), dtype=torch.int32, device=dev)
            a = _abi.PfRasterVisArgs()
            a.cam, ray_maps = self._camera(camera)
            a.scene = _scene_struct(all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency,
                                    adjacency_offsets, adjacency_diff)
            a.tiles = self._tiles(nt, offsets, tile_sites, order)
            a.transmittance_threshold = float(vis_options.transmittance_threshold)
            a.max_intersections = int(vis_options.max_intersections)
            a.depth_quantile = float(vis_options.depth_quantile)
            a.bkgd_color[:] = [float(v) for v in vis_options.bkgd_color]
            a.color, a.depth, a.normal = color.data_ptr(), depth.data_ptr(), normal.data_ptr()
            a.alpha, a.intersections = alpha.data_ptr(), inter.data_ptr()
            self._launch("raster_vis", self.lib.pf_raster_vis, _ref(a), stream_handle())
            return color, depth, normal, alpha, inter

    # ----------------------------------------------------------------- backward
    def _backward(self, grad_ctx, grad_color_in, grad_opacity_in, grad_normal_distance_in, grad_normal_in,
                  grad_quantile_depths_in, grad_err_in, grad_contrib_in, grad_point_err_in,
                  grad_prim_visible_mask_in):
        del grad_point_err_in, grad_prim_visible_mask_in  # unused, as the reference
        if self.attr_dtype != "float":
            raise ValueError("Backward pass only supports float precision")
        if self.loss_accumulation:
            raise NotImplementedError()  # rasterize.py:1232-1234
        (dq, all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency, adjacency_offsets,
         adjacency_diff, tile_sites, offsets, tile_early_stop, ray_gt, color_out, log_t_out) = grad_ctx.saved_tensors
        camera = grad_ctx.camera
        H, W = camera.height, camera.width
        N = all_spheres.shape[0]
        T = all_texel_sites.shape[1]
        dev = all_spheres.device
        # the five gradients are views of one zeroed buffer (one fill, one
        # scrub); its tail, 256-byte aligned, stands in for every incoming
        # gradient that is None (the kernel only reads those)
        shapes = ((N, 4), (N, 4), (N, T, 3), (N, T, 3), (N, T))
        sizes = [int(np.prod(sh)) for sh in shapes]
        n_grads = sum(sizes)
        zero_at = -(-n_grads // 64) * 64
        missing = [(H * W * 3, grad_color_in), (H * W, grad_opacity_in), (H * W, grad_normal_distance_in),
                   (H * W * 3, grad_normal_in), (N, grad_contrib_in)]
        if dq is not None:
            missing.append((int(np.prod(dq.shape)), grad_quantile_depths_in))
        n_zero = max([n for n, t in missing if t is None], default=0)
        grad_buf = torch.zeros(zero_at + n_zero, dtype=torch.float32, device=dev)
        spheres_grad, nsigmas_grad, texel_sites_grad, texel_rgb_grad, texel_height_grad = (
            v.view(sh) for v, sh in zip(torch.split(grad_buf[:n_grads], sizes), shapes))
        g = lambda t, shape: _f32c(t) if t is not None else grad_buf[zero_at: zero_at + int(np.prod(shape))].view(shape)
        grad_color = g(grad_color_in, (H, W, 3))
        if grad_opacity_in is not None:
            grad_log_t = (-_f32c(grad_opacity_in) * torch.exp(log_t_out)).contiguous()
        else:
            grad_log_t = g(None, (H, W))
        grad_nd = g(grad_normal_distance_in, (H, W))
        grad_n = g(grad_normal_in, (H, W, 3))
        grad_contrib = g(grad_contrib_in, (N,))
        grad_q = None
        if dq is not None:
            grad_q = g(grad_quantile_depths_in, tuple(dq.shape))

        a = _abi.PfRasterBwdArgs()
        a.cam, ray_maps = self._camera(camera, grad_ctx.ray_maps)
        a.scene = _scene_struct(all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency,
                                adjacency_offsets, None if self.exact_geometry else adjacency_diff,
                                getattr(grad_ctx, "rgbh_bf16", None), grad_ctx.face_count)
        sg = grad_ctx.segs if self.split_backward else None
        order = grad_ctx.order
        if sg is None and order is None:
            order = self.tile_order(offsets)
        a.tiles = self._tiles(offsets.shape[0] - 1, offsets, tile_sites, order)
        a.segs = self._segs(sg, torch.empty(sg["n_segs"] * 64, dtype=torch.float32, device=dev)
                            if sg is not None else None)
        a.depth_quantiles = dq.data_ptr() if dq is not None else None
        a.n_quantiles = dq.shape[-1] if dq is not None else 0
        a.tile_early_stop = tile_early_stop.data_ptr()
        a.log_t = log_t_out.data_ptr()
        a.grad_color = grad_color.data_ptr()
        a.grad_log_t = grad_log_t.data_ptr()
        a.grad_normal_distance = grad_nd.data_ptr()
        a.grad_normal = grad_n.data_ptr()
        a.grad_quantile_depths = grad_q.data_ptr() if grad_q is not None else None
        a.grad_contrib = grad_contrib.data_ptr()
        a.exact_gradients = int(bool(self.exact_gradients))
        es = getattr(grad_ctx, "entry_shades", None)
        a.entry_shades = es.data_ptr() if es is not None else None
        a.grad_spheres = spheres_grad.data_ptr()
        a.grad_nsigmas = nsigmas_grad.data_ptr()
        a.grad_texel_sites = texel_sites_grad.data_ptr()
        a.grad_texel_rgb = texel_rgb_grad.data_ptr()
        a.grad_texel_height = texel_height_grad.data_ptr()
        st = stream_handle()
        self._launch("raster_bwd", self.lib.pf_raster_bwd, _ref(a), st)
        check(self.lib.pf_scrub_nonfinite(ptr(grad_buf), n_grads, st))
        head = grad_ctx.color_head
        del grad_ctx.rasterizer, grad_ctx.camera, grad_ctx.ray_maps, grad_ctx.order, grad_ctx.segs, grad_ctx.rgbh_bf16
        head_sites = grad_ctx.head_sites
        del grad_ctx.color_head, grad_ctx.face_count, grad_ctx.head_sites
        grads = (None, None, None, spheres_grad[:, :3], spheres_grad[:, 3], nsigmas_grad[:, 3],
                 nsigmas_grad[:, :3], texel_sites_grad, texel_rgb_grad, texel_height_grad, None, None, None, None)
        if head is None:
            return grads
        # the head's backward over the texel colour gradients (a texel's
        # gradient is summed over every ray and tile first, then taken
        # through the network once)
        eye = torch.as_tensor(camera.eye).detach().to("cpu", torch.float32).reshape(-1)
        rows = dict(n_rows=N * T, T=T, texel_sites=all_texel_sites, eye=tuple(float(v) for v in eye))
        head_grads = head._run_bwd(rows, texel_rgb_grad.view(N * T, 3), site_list=head_sites)
        return grads[:8] + (None,) + grads[9:] + (None, None) + tuple(head_grads)


def power_face_encoding(points, radii, adjacency, adjacency_offsets):
    """The benchmark's fp16 face encoding (benchmark.py:297-302): per edge
    (c_adj - c, pm_adj - pm) with pm = (|c|^2 - r^2)/2, computed in torch as the
    reference does."""
    num_adjs = adjacency_offsets[1:] - adjacency_offsets[:-1]
    pm = 0.5 * (points.norm(dim=-1) ** 2 - radii ** 2)
    adj = adjacency.long()
    self_points = points.repeat_interleave(num_adjs.long(), dim=0)
    diff = points[adj] - self_points
    pm_diff = pm[adj] - pm.repeat_interleave(num_adjs.long(), dim=0)
    return torch.cat([diff, pm_diff[:, None]], dim=-1).to(torch.float16)


def _ref(struct):
    return ctypes.byref(struct)


def _head_encoding(head):
    from .mlp_head import ENCODINGS
    return ENCODINGS[head.encoding]


def _scene_struct(all_spheres, all_nsigmas, all_texel_sites, all_texel_rgbh, adjacency, adjacency_offsets,
                  adjacency_diff, texel_rgbh_bf16=None, adjacency_count=None):
    s = _abi.PfScene()
    s.n_sites = all_spheres.shape[0]
    s.n_texels = all_texel_sites.shape[1]
    s.spheres = all_spheres.data_ptr()
    s.nsigmas = all_nsigmas.data_ptr()
    s.texel_sites = all_texel_sites.data_ptr()
    s.texel_rgbh = all_texel_rgbh.data_ptr()
    s.adjacency = adjacency.data_ptr() if adjacency.numel() else None
    s.adjacency_offsets = adjacency_offsets.data_ptr()
    s.adjacency_diff = adjacency_diff.data_ptr() if (adjacency_diff is not None and adjacency_diff.numel()) else None
    s.texel_rgbh_bf16 = texel_rgbh_bf16.data_ptr() if texel_rgbh_bf16 is not None else None
    s.adjacency_count = adjacency_count.data_ptr() if adjacency_count is not None else None
    return s
