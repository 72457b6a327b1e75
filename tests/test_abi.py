"""The C-ABI library loads on a CPU-only host and exports every entry point
include/powerfoam_amd.h declares; the ctypes struct mirror matches the C
layout (checked against the gcc-built oracle, which includes the header)."""

import ctypes
import os
import re

import pytest

from powerfoam_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_declared_symbols():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    lib = ctypes.CDLL(_abi.LIB_PATH)
    decl = declared_symbols()
    assert {"pf_raster_fwd", "pf_raster_bwd", "pf_bin_count", "pf_bin_sort"} <= decl
    for name in sorted(decl):
        assert hasattr(lib, name), name
    # every declared symbol has a ctypes signature and vice versa
    assert decl == set(_abi.SIGNATURES)


def test_load_sets_signatures_and_reports_errors():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    lib = _abi.load()
    assert lib.pf_version() >= 1
    # argument validation runs on the host, before any device call
    st = lib.pf_bin_scan(None, 0, None, None)
    assert st == _abi.PF_ERR_ARG
    assert b"pf_bin_scan" in lib.pf_last_error()
    with pytest.raises(_abi.PfError):
        _abi.check(st)


def test_struct_layout_matches_c():
    """ctypes mirror vs gcc's layout of the same header (the oracle library
    includes include/powerfoam_amd.h and reports sizeof/offsetof)."""
    import numpy as np
    from oracle import orc
    lib = orc.load()
    lib.orc_abi_layout.argtypes = [ctypes.c_void_p, ctypes.c_int]
    out = np.zeros(64, np.int64)
    n = lib.orc_abi_layout(out.ctypes.data, 64)
    C, S, TL, F, B = _abi.PfCamera, _abi.PfScene, _abi.PfTileLists, _abi.PfRasterFwdArgs, _abi.PfRasterBwdArgs
    RG, V, TA, TB = _abi.PfRasterRgbArgs, _abi.PfRasterVisArgs, _abi.PfTraceArgs, _abi.PfTraceBwdArgs
    MR, SG = _abi.PfMlpRows, _abi.PfSegments
    expect = [ctypes.sizeof(C), C.ray_maps.offset, C.fov_cos_cutoff.offset,
              ctypes.sizeof(S), S.adjacency_diff.offset, S.adjacency_count.offset,
              ctypes.sizeof(TL), TL.sites.offset,
              ctypes.sizeof(F), F.tiles.offset, F.transmittance_threshold.offset, F.point_err.offset,
              F.entry_shades.offset, F.mlp_head.offset,
              ctypes.sizeof(_abi.PfMlpHead), _abi.PfMlpHead.latents.offset, _abi.PfMlpHead.texel_rgbh_out.offset,
              ctypes.sizeof(B), B.exact_gradients.offset, B.grad_texel_height.offset, B.entry_shades.offset,
              ctypes.sizeof(RG), RG.color.offset,
              ctypes.sizeof(V), V.bkgd_color.offset, V.intersections.offset,
              ctypes.sizeof(TA), TA.max_steps.offset, TA.n_cells.offset, TA.tile_steps.offset, TA.face_planes.offset,
              ctypes.sizeof(TB), TB.grad_opacity.offset, TB.grad_texel_height.offset, TB.tile_order.offset,
              ctypes.sizeof(MR), MR.eye.offset, MR.latents.offset, MR.site_count.offset, MR.out_stride.offset,
              ctypes.sizeof(_abi.PfMlpGrads),
              ctypes.sizeof(SG), SG.count.offset, SG.seg_len.offset, SG.scratch.offset]
    assert n == len(expect)
    assert out[:n].tolist() == expect


def test_live_sites_argument_checks_run_on_the_host():
    """pf_live_sites refuses unknown face encodings and ray-map cameras before
    any device work (the addresses below are never dereferenced)."""
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    lib = _abi.load()
    cam = _abi.PfCamera()
    cam.width, cam.height = 16, 16
    fake = ctypes.c_void_p(0x1000)
    st = lib.pf_live_sites(ctypes.byref(cam), fake, 4, fake, fake, 10, fake, 5, fake, None, None)
    assert st == _abi.PF_ERR_ARG and b"encoding" in lib.pf_last_error()
    st = lib.pf_live_sites(ctypes.byref(cam), fake, 4, None, fake, 10, None, 0, fake, None, None)
    assert st == _abi.PF_ERR_ARG
    cam.ray_maps = 0x2000
    st = lib.pf_live_sites(ctypes.byref(cam), fake, 4, fake, fake, 10, None, 0, fake, None, None)
    assert st == _abi.PF_ERR_UNSUPPORTED


def test_adjacency_encode_rejects_null_edges_on_the_host():
    """pf_adj_encode with edges but no adjacency or output is an argument
    error before any device work (no GPU needed)."""
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    lib = _abi.load()
    assert lib.pf_adj_encode(1, 10, None, 1, 5, 1, None) == _abi.PF_ERR_ARG
    assert b"NULL adjacency" in lib.pf_last_error()
    assert lib.pf_adj_encode(1, 10, 1, 1, 5, None, None) == _abi.PF_ERR_ARG
    assert lib.pf_adj_encode(1, 10, 1, 1, -1, 1, None) == _abi.PF_ERR_ARG
    # no edges: nothing is read, nothing to reject
    assert lib.pf_adj_encode(1, 10, None, 1, 0, None, None) == _abi.PF_OK


def test_trace_tile_order_argument_checks_run_on_the_host():
    """pf_trace_tile_order: nothing to order is not an error, NULL arrays or a
    negative count are (no device work either way)."""
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    lib = _abi.load()
    assert lib.pf_trace_tile_order(None, 0, None, None) == _abi.PF_OK
    assert lib.pf_trace_tile_order(None, 4, None, None) == _abi.PF_ERR_ARG
    assert lib.pf_trace_tile_order(None, -1, None, None) == _abi.PF_ERR_ARG


def test_trace_planes_argument_checks_run_on_the_host():
    """pf_trace_planes: no sites is not an error; a missing scene array or
    output is, before any device work."""
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    lib = _abi.load()
    sc = _abi.PfScene()
    sc.n_sites = 0
    assert lib.pf_trace_planes(ctypes.byref(sc), 5, None, None) == _abi.PF_OK
    sc.n_sites = 10
    assert lib.pf_trace_planes(ctypes.byref(sc), 0, None, None) == _abi.PF_OK
    assert lib.pf_trace_planes(ctypes.byref(sc), 5, None, None) == _abi.PF_ERR_ARG
    assert lib.pf_trace_planes(ctypes.byref(sc), -1, None, None) == _abi.PF_ERR_ARG
    assert lib.pf_trace_planes(None, 5, None, None) == _abi.PF_ERR_ARG


def test_site_list_argument_checks_run_on_the_host():
    """pf_mlp_active_sites takes 1..8 texels per site and pf_mask_sites needs
    its mask and outputs; no sites is not an error (no device work either
    way)."""
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    lib = _abi.load()
    for t in (0, 9):
        assert lib.pf_mlp_active_sites(1, 10, t, 1, 1, None) == _abi.PF_ERR_ARG
    assert lib.pf_mlp_active_sites(None, 10, 8, 1, 1, None) == _abi.PF_ERR_ARG
    assert lib.pf_mlp_active_sites(None, 0, 8, None, None, None) == _abi.PF_OK
    assert lib.pf_mask_sites(None, None, 10, 1, 1, None) == _abi.PF_ERR_ARG
    assert lib.pf_mask_sites(1, None, -1, 1, 1, None) == _abi.PF_ERR_ARG
    assert lib.pf_mask_sites(None, None, 0, None, None, None) == _abi.PF_OK
