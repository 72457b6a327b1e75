"""Cech adjacency on the GPU, behind the reference's interface
(powerfoam/bvh.py:72-123: AABBTree(device), .update(centers, radii),
.build_cech_complex() -> (adjacent, offsets), both int32).

The neighbour set is the reference's (count_adjacent / write_adjacent,
bvh.py:7-69): j != i when the float32 boxes c -+ r overlap and
|c_i - c_j| < r_i + r_j with c, r taken back from the box. The search is a
hashed grid per radius level (pf_cech_count / pf_cech_emit) in place of
Warp's BVH, and each site's ids come out ascending (the BVH's order is
unspecified). Like the reference, one host sync reads the edge total before
the lists are written.
"""

import torch

from . import _abi
from ._abi import PfError, check, ptr, stream_handle


class AABBTree:
    def __init__(self, device):
        self.device = torch.device(device)
        self.lib = _abi.load()
        self.spheres = None

    def update(self, centers, radii):
        """Takes the sites for the next build (bvh.py:76-94); min / max are the
        float32 boxes the reference keeps."""
        if centers.ndim != 2 or centers.shape[1] != 3 or radii.shape != centers.shape[:1]:
            raise ValueError("centers must be [N,3] and radii [N]")
        if centers.shape[0] > _abi.CECH_MAX_SITES:
            raise PfError(f"{centers.shape[0]} sites exceed the Cech builder's {_abi.CECH_MAX_SITES}")
        c = centers.detach().to(self.device, torch.float32)
        r = radii.detach().to(self.device, torch.float32)
        self.spheres = torch.cat([c, r[:, None]], 1).contiguous()
        self.min = c - r[:, None]
        self.max = c + r[:, None]

    def build_cech_complex(self):
        """(adjacent int32 [M], offsets int32 [N+1]) -- bvh.py:96-123."""
        if self.spheres is None:
            raise PfError("build_cech_complex before update")
        lib, dev = self.lib, self.device
        n = self.spheres.shape[0]
        st = stream_handle()
        ws_bytes = int(lib.pf_cech_workspace_size(n))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        counts = torch.empty(n + 1, dtype=torch.int32, device=dev)
        check(lib.pf_cech_count(ptr(self.spheres), n, ptr(ws), ws_bytes, ptr(counts), st))
        offsets = torch.cumsum(counts, 0)  # int64
        m = int(offsets[-1])
        if m >= 2 ** 31:
            raise PfError(f"Cech adjacency of {m} entries overflows int32 offsets")
        adjacent = torch.empty(m, dtype=torch.int32, device=dev)
        if m > 0:
            records = torch.empty(m, dtype=torch.int64, device=dev)
            dropped = torch.zeros(1, dtype=torch.int32, device=dev)
            check(lib.pf_cech_emit(ptr(self.spheres), n, ptr(ws), ws_bytes, ptr(offsets), ptr(records),
                                   ptr(dropped), st))
            del ws
            sort_bytes = int(lib.pf_bin_sort_workspace_size(m, n))
            sort_ws = torch.empty(max(sort_bytes, 1), dtype=torch.uint8, device=dev)
            check(lib.pf_bin_sort(ptr(offsets), n, ptr(records), ptr(adjacent), ptr(sort_ws), sort_bytes, st))
            if int(dropped) != 0:
                raise PfError(f"Cech emit dropped {int(dropped)} entries that the count pass did not see")
        adjacent._pf_pairing = 2  # symmetric lists (read by InterpenetrationFunction)
        return adjacent, offsets.to(torch.int32)


def restricted_power_adjacency(centers, radii, adjacency=None, offsets=None):
    """The power faces that bound each site's cell inside its ball
    (pf_power_restrict): from the Cech lists (built here when not given), the
    entries whose radical-plane face meets the two balls, plus one-way faces
    for sites whose cell misses their ball. Renders as the Cech adjacency does,
    with the faces that lie outside the balls left out. Returns (adjacency
    int32, offsets int32)."""
    dev = centers.device
    if adjacency is None:
        tree = AABBTree(dev)
        tree.update(centers, radii)
        adjacency, offsets = tree.build_cech_complex()
    lib = _abi.load()
    n = centers.shape[0]
    if offsets.shape != (n + 1,):
        raise ValueError("offsets must be [N+1]")
    sp = torch.cat([centers.detach().float(), radii.detach().float()[:, None]], 1).contiguous()
    adjacency = adjacency.to(dev, torch.int32).contiguous()
    offsets = offsets.to(dev, torch.int32).contiguous()
    m = adjacency.numel()
    keep = torch.empty(max(m, 1), dtype=torch.uint8, device=dev)
    counts = torch.empty(n + 1, dtype=torch.int32, device=dev)
    check(lib.pf_power_restrict(ptr(sp), n, ptr(adjacency), ptr(offsets), m, ptr(keep), ptr(counts),
                                stream_handle()))
    out = adjacency[keep[:m].bool()]
    out._pf_pairing = 1  # ascending without repeats; one-way entries for sites outside their cells
    return out, torch.cumsum(counts, 0).to(torch.int32)
