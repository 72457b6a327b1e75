"""Morton site ordering and the interpenetration regulariser on the GPU,
behind the reference's names (powerfoam/geometry.py:50-189): morton_sort(points)
returns the int32 order of the points along the Z-order curve of their
bounding box, InterpenetrationFunction.apply(points, radii, adjacency,
adjacency_offsets) the per-site sum of squared overlaps with its neighbours,
differentiable w.r.t. points and radii.
"""

import torch

from . import _abi
from ._abi import check, ptr, stream_handle


def morton_codes_and_order(points):
    """(codes int64 [N] per point, order int32 [N]) -- pf_morton_sort."""
    lib = _abi.load()
    n = points.shape[0]
    dev = points.device
    p = points.detach().to(torch.float32).contiguous()
    codes = torch.empty(n, dtype=torch.int64, device=dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    if n == 0:
        return codes, order
    ws_bytes = int(lib.pf_morton_workspace_size(n))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    check(lib.pf_morton_sort(ptr(p), n, ptr(ws), ws_bytes, ptr(codes), ptr(order), stream_handle()))
    return codes, order


def morton_sort(points):
    """Sorted indices of the points along the Morton curve (geometry.py:50-85)."""
    return morton_codes_and_order(points)[1]


class InterpenetrationFunction(torch.autograd.Function):
    """geometry.py:120-189: areas[i] = sum_j (r_i + r_j - |c_i - c_j|)^2 over
    i's overlapping neighbours. `pairing` (an addition; default: what the
    list builder recorded on the adjacency tensor, else 0) lets the backward
    sum the entries whose reverse is in the other list without atomics:
    2 for symmetric lists (AABBTree.build_cech_complex), 1 for ascending lists
    without repeats (restricted_power_adjacency)."""

    @staticmethod
    def forward(ctx, points, radii, adjacency, adjacency_offsets, pairing=None):
        lib = _abi.load()
        n = points.shape[0]
        ctx.pairing = int(getattr(adjacency, "_pf_pairing", 0) if pairing is None else pairing)
        spheres = torch.cat([points.detach(), radii.detach()[:, None]], -1).to(torch.float32).contiguous()
        adjacency = adjacency.to(torch.int32).contiguous()
        adjacency_offsets = adjacency_offsets.to(torch.int32).contiguous()
        if adjacency_offsets.shape != (n + 1,):
            raise ValueError("adjacency_offsets must be [N+1]")
        areas = torch.empty(n, dtype=torch.float32, device=points.device)
        check(lib.pf_interpenetration_fwd(ptr(spheres), n, ptr(adjacency), ptr(adjacency_offsets),
                                          adjacency.numel(), ptr(areas), stream_handle()))
        ctx.save_for_backward(spheres, adjacency, adjacency_offsets)
        return areas

    @staticmethod
    def backward(ctx, grad_areas):
        lib = _abi.load()
        spheres, adjacency, adjacency_offsets = ctx.saved_tensors
        n = spheres.shape[0]
        grad = torch.zeros_like(spheres)
        g = grad_areas.to(torch.float32).contiguous()
        check(lib.pf_interpenetration_bwd(ptr(spheres), n, ptr(adjacency), ptr(adjacency_offsets),
                                          adjacency.numel(), ctx.pairing, ptr(g), ptr(grad), stream_handle()))
        return grad[:, :3], grad[:, 3], None, None, None
