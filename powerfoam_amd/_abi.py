"""ctypes mirror of include/powerfoam_amd.h and the loader of libpfhip.so.

This module replaces the reference's Warp launch shim (powerfoam/warp_interop.py:6-24):
every kernel is reached through the C-ABI shared library built from
powerfoam_amd/csrc, with raw device pointers and torch's current HIP stream.
There is no fallback: if the library is missing, loading raises.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpfhip.so")

PF_OK = 0
PF_ERR_ARG = -1
PF_ERR_OVERFLOW = -2
PF_ERR_HIP = -3
PF_ERR_UNSUPPORTED = -4
PF_FACES_RADIUS = 0  # (c_adj - c, r_adj) from pf_adj_encode
PF_FACES_POWER = 1  # (c_adj - c, pm_adj - pm), the benchmark's encoding

TILE_WIDTH = 8
TILE_SIZE = 64
MAX_TEXELS = 8
SORT_STATIC_MAX_TILES = 262144  # PF_SORT_STATIC_MAX_TILES
CECH_MAX_SITES = 1 << 29  # PF_CECH_MAX_SITES
MLP_ENC_SH = 0  # PF_MLP_ENC_SH
MLP_ENC_FREQ = 1  # PF_MLP_ENC_FREQ

c_float_p = ctypes.POINTER(ctypes.c_float)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
c_i64_p = ctypes.POINTER(ctypes.c_int64)
c_u16_p = ctypes.POINTER(ctypes.c_uint16)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)


class PfCamera(ctypes.Structure):
    _fields_ = [
        ("eye", ctypes.c_float * 3),
        ("right", ctypes.c_float * 3),
        ("up", ctypes.c_float * 3),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("ray_maps", ctypes.c_void_p),
        ("fov_cos_cutoff", ctypes.c_float),
    ]


class PfScene(ctypes.Structure):
    _fields_ = [
        ("n_sites", ctypes.c_int32),
        ("n_texels", ctypes.c_int32),
        ("spheres", ctypes.c_void_p),
        ("nsigmas", ctypes.c_void_p),
        ("texel_sites", ctypes.c_void_p),
        ("texel_rgbh", ctypes.c_void_p),
        ("adjacency", ctypes.c_void_p),
        ("adjacency_offsets", ctypes.c_void_p),
        ("adjacency_diff", ctypes.c_void_p),
        ("texel_rgbh_bf16", ctypes.c_void_p),
        ("adjacency_count", ctypes.c_void_p),
    ]


class PfSegments(ctypes.Structure):
    _fields_ = [
        ("n_segs", ctypes.c_int32),
        ("count", ctypes.c_void_p),
        ("seg_len", ctypes.c_int32),
        ("tile_first", ctypes.c_void_p),
        ("seg_tile", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("log_t_sums", ctypes.c_void_p),
        ("lane_active", ctypes.c_void_p),
        ("scratch", ctypes.c_void_p),
    ]


class PfTileLists(ctypes.Structure):
    _fields_ = [
        ("n_tiles", ctypes.c_int32),
        ("offsets", ctypes.c_void_p),
        ("sites", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
    ]


class PfRasterFwdArgs(ctypes.Structure):
    _fields_ = [
        ("cam", PfCamera),
        ("scene", PfScene),
        ("tiles", PfTileLists),
        ("segs", PfSegments),
        ("depth_quantiles", ctypes.c_void_p),
        ("n_quantiles", ctypes.c_int32),
        ("ray_gt", ctypes.c_void_p),
        ("transmittance_threshold", ctypes.c_float),
        ("return_point_err", ctypes.c_int32),
        ("loss_accumulation", ctypes.c_int32),
        ("tile_early_stop", ctypes.c_void_p),
        ("color", ctypes.c_void_p),
        ("log_t", ctypes.c_void_p),
        ("normal_distance", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("quantile_depths", ctypes.c_void_p),
        ("err", ctypes.c_void_p),
        ("contrib", ctypes.c_void_p),
        ("point_err", ctypes.c_void_p),
        ("entry_shades", ctypes.c_void_p),
        ("mlp_head", ctypes.c_void_p),
    ]


class PfRasterBwdArgs(ctypes.Structure):
    _fields_ = [
        ("cam", PfCamera),
        ("scene", PfScene),
        ("tiles", PfTileLists),
        ("segs", PfSegments),
        ("depth_quantiles", ctypes.c_void_p),
        ("n_quantiles", ctypes.c_int32),
        ("tile_early_stop", ctypes.c_void_p),
        ("log_t", ctypes.c_void_p),
        ("grad_color", ctypes.c_void_p),
        ("grad_log_t", ctypes.c_void_p),
        ("grad_normal_distance", ctypes.c_void_p),
        ("grad_normal", ctypes.c_void_p),
        ("grad_quantile_depths", ctypes.c_void_p),
        ("grad_contrib", ctypes.c_void_p),
        ("exact_gradients", ctypes.c_int32),
        ("grad_spheres", ctypes.c_void_p),
        ("grad_nsigmas", ctypes.c_void_p),
        ("grad_texel_sites", ctypes.c_void_p),
        ("grad_texel_rgb", ctypes.c_void_p),
        ("grad_texel_height", ctypes.c_void_p),
        ("entry_shades", ctypes.c_void_p),
    ]


class PfRasterRgbArgs(ctypes.Structure):
    _fields_ = [
        ("cam", PfCamera),
        ("scene", PfScene),
        ("tiles", PfTileLists),
        ("transmittance_threshold", ctypes.c_float),
        ("color", ctypes.c_void_p),
    ]


class PfRasterVisArgs(ctypes.Structure):
    _fields_ = [
        ("cam", PfCamera),
        ("scene", PfScene),
        ("tiles", PfTileLists),
        ("transmittance_threshold", ctypes.c_float),
        ("max_intersections", ctypes.c_int32),
        ("depth_quantile", ctypes.c_float),
        ("bkgd_color", ctypes.c_float * 3),
        ("color", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("alpha", ctypes.c_void_p),
        ("intersections", ctypes.c_void_p),
    ]


class PfMlpWeights(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("w1", "b1", "w2", "b2", "w3", "b3")]


class PfMlpHead(ctypes.Structure):
    _fields_ = [
        ("weights", PfMlpWeights),
        ("latents", ctypes.c_void_p),
        ("encoding", ctypes.c_int32),
        ("texel_rgbh_out", ctypes.c_void_p),
    ]


class PfMlpRows(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_int32),
        ("n_sites", ctypes.c_int32),
        ("n_texels", ctypes.c_int32),
        ("texel_sites", ctypes.c_void_p),
        ("eye", ctypes.c_float * 3),
        ("pair_sites", ctypes.c_void_p),
        ("pair_dirs", ctypes.c_void_p),
        ("latents", ctypes.c_void_p),
        ("encoding", ctypes.c_int32),
        ("site_list", ctypes.c_void_p),
        ("site_count", ctypes.c_void_p),
        ("out_stride", ctypes.c_int32),
    ]


class PfMlpGrads(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("latents", "w1", "b1", "w2", "b2", "w3", "b3")]


class PfTraceArgs(ctypes.Structure):
    _fields_ = [
        ("cam", PfCamera),
        ("scene", PfScene),
        ("start_site", ctypes.c_int32),
        ("start_sites", ctypes.c_void_p),
        ("transmittance_threshold", ctypes.c_float),
        ("depth_quantile", ctypes.c_float),
        ("max_steps", ctypes.c_int32),
        ("color", ctypes.c_void_p),
        ("log_t", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("q_step", ctypes.c_void_p),
        ("q_sigma", ctypes.c_void_p),
        ("cells", ctypes.c_void_p),
        ("max_cells", ctypes.c_int32),
        ("n_cells", ctypes.c_void_p),
        ("compact", ctypes.c_int32),
        ("tile_steps", ctypes.c_void_p),
        ("face_planes", ctypes.c_void_p),
    ]


class PfTraceBwdArgs(ctypes.Structure):
    _fields_ = [("fwd", PfTraceArgs)] + [(k, ctypes.c_void_p) for k in (
        "color", "normal", "log_t", "q_step", "q_sigma", "grad_color", "grad_normal", "grad_depth",
        "grad_opacity", "grad_spheres", "grad_nsigmas", "grad_texel_sites", "grad_texel_rgb",
        "grad_texel_height", "tile_order")]


class PfTraceCheckArgs(ctypes.Structure):
    _fields_ = [("cam", PfCamera), ("scene", PfScene), ("n_rays", ctypes.c_int32), ("pixels", ctypes.c_void_p),
                ("cells", ctypes.c_void_p), ("n_cells", ctypes.c_void_p), ("max_cells", ctypes.c_int32),
                ("t_enter", ctypes.c_void_p), ("t_exit", ctypes.c_void_p), ("next_cell", ctypes.c_void_p),
                ("exit_gap", ctypes.c_void_p), ("exit_tol", ctypes.c_void_p)]


class PfSvArgs(ctypes.Structure):
    _fields_ = [
        ("n_points", ctypes.c_int32),
        ("sv_dof", ctypes.c_int32),
        ("points", ctypes.c_void_p),
        ("eye", ctypes.c_float * 3),
        ("forward", ctypes.c_float * 3),
        ("fov_cos_cutoff", ctypes.c_float),
        ("att_sites", ctypes.c_void_p),
        ("att_values", ctypes.c_void_p),
        ("att_temps", ctypes.c_void_p),
    ]


class PfGlueParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("points", "radii", "density", "quaternions", "texel_offsets",
                                               "texel_height")]


class PfGlueOutputs(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("radii", "density", "normals", "texel_sites", "texel_height")]


class PfGlueGrads(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("points", "radii", "density", "quaternions", "texel_offsets",
                                               "texel_height")]


# Exported C symbols and their ctypes signatures (argtypes, restype).
_VP = ctypes.c_void_p
SIGNATURES = {
    "pf_version": ([], ctypes.c_int),
    "pf_last_error": ([], ctypes.c_char_p),
    "pf_bin_count": ([_VP, _VP, _VP, _VP, ctypes.c_int32, _VP, _VP], ctypes.c_int),
    "pf_bin_scan": ([_VP, ctypes.c_int32, _VP, _VP], ctypes.c_int),
    "pf_bin_emit": ([_VP, _VP, _VP, _VP, ctypes.c_int32, _VP, ctypes.c_int32, _VP, _VP, _VP, _VP, _VP],
                    ctypes.c_int),
    "pf_live_sites": ([_VP, _VP, ctypes.c_int32, _VP, _VP, ctypes.c_int64, _VP, ctypes.c_int32, _VP, _VP, _VP],
                      ctypes.c_int),
    "pf_bin_sort_workspace_size": ([ctypes.c_int64, ctypes.c_int32], ctypes.c_size_t),
    "pf_bin_sort": ([_VP, ctypes.c_int32, _VP, _VP, _VP, ctypes.c_size_t, _VP], ctypes.c_int),
    "pf_bin_sort_static": ([_VP, ctypes.c_int32, _VP, _VP, _VP], ctypes.c_int),
    "pf_tile_schedule": ([_VP, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _VP, _VP, _VP, _VP, _VP,
                          _VP], ctypes.c_int),
    "pf_cone_tree_nodes": ([ctypes.c_int32, ctypes.c_int32], ctypes.c_int32),
    "pf_bin_cones": ([_VP, _VP, _VP], ctypes.c_int),
    "pf_bin_count_cones": ([_VP, _VP, _VP, ctypes.c_int32, _VP, _VP], ctypes.c_int),
    "pf_bin_emit_cones": ([_VP, _VP, _VP, ctypes.c_int32, _VP, ctypes.c_int32, _VP, _VP, _VP, _VP, _VP],
                          ctypes.c_int),
    "pf_adj_encode": ([_VP, ctypes.c_int32, _VP, _VP, ctypes.c_int64, _VP, _VP], ctypes.c_int),
    "pf_cech_workspace_size": ([ctypes.c_int32], ctypes.c_size_t),
    "pf_cech_count": ([_VP, ctypes.c_int32, _VP, ctypes.c_size_t, _VP, _VP], ctypes.c_int),
    "pf_cech_emit": ([_VP, ctypes.c_int32, _VP, ctypes.c_size_t, _VP, _VP, _VP, _VP], ctypes.c_int),
    "pf_prune_faces": ([_VP, ctypes.c_int32, _VP, _VP, ctypes.c_int64, _VP, ctypes.c_float, _VP, _VP, _VP, _VP],
                       ctypes.c_int),
    "pf_power_restrict": ([_VP, ctypes.c_int32, _VP, _VP, ctypes.c_int64, _VP, _VP, _VP], ctypes.c_int),
    "pf_morton_workspace_size": ([ctypes.c_int32], ctypes.c_size_t),
    "pf_morton_sort": ([_VP, ctypes.c_int32, _VP, ctypes.c_size_t, _VP, _VP, _VP], ctypes.c_int),
    "pf_interpenetration_fwd": ([_VP, ctypes.c_int32, _VP, _VP, ctypes.c_int64, _VP, _VP], ctypes.c_int),
    "pf_interpenetration_bwd": ([_VP, ctypes.c_int32, _VP, _VP, ctypes.c_int64, ctypes.c_int32, _VP, _VP, _VP],
                                ctypes.c_int),
    "pf_raster_fwd": ([_VP, _VP], ctypes.c_int),
    "pf_raster_bwd": ([_VP, _VP], ctypes.c_int),
    "pf_raster_rgb": ([_VP, _VP], ctypes.c_int),
    "pf_raster_vis": ([_VP, _VP], ctypes.c_int),
    "pf_mlp_fwd": ([_VP, _VP, _VP, _VP], ctypes.c_int),
    "pf_mlp_bwd": ([_VP, _VP, _VP, _VP, _VP], ctypes.c_int),
    "pf_mlp_active_sites": ([_VP, ctypes.c_int32, ctypes.c_int32, _VP, _VP, _VP], ctypes.c_int),
    "pf_mask_sites": ([_VP, _VP, ctypes.c_int32, _VP, _VP, _VP], ctypes.c_int),
    "pf_trace_fwd": ([_VP, _VP], ctypes.c_int),
    "pf_trace_bwd": ([_VP, _VP], ctypes.c_int),
    "pf_trace_tile_order": ([_VP, ctypes.c_int32, _VP, _VP], ctypes.c_int),
    "pf_trace_planes": ([_VP, ctypes.c_int64, _VP, _VP], ctypes.c_int),
    "pf_trace_check_f64": ([_VP, _VP], ctypes.c_int),
    "pf_pack_sites": ([_VP] * 6 + [ctypes.c_int32, ctypes.c_int32] + [_VP] * 4, ctypes.c_int),
    "pf_scrub_nonfinite": ([_VP, ctypes.c_int64, _VP], ctypes.c_int),
    "pf_probe_valu": ([ctypes.c_int32, ctypes.c_int32, _VP, _VP], ctypes.c_int64),
    "pf_glue_fwd": ([_VP, ctypes.c_int32, ctypes.c_int32, _VP, _VP], ctypes.c_int),
    "pf_sv_fwd": ([_VP, _VP, _VP, _VP], ctypes.c_int),
    "pf_sv_bwd": ([_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP], ctypes.c_int),
    "pf_glue_bwd": ([_VP, ctypes.c_int32, ctypes.c_int32, _VP, _VP, _VP], ctypes.c_int),
    "pf_sh_fwd": ([_VP, _VP, ctypes.c_int32, ctypes.c_int32, _VP, _VP, _VP], ctypes.c_int),
    "pf_sh_bwd": ([_VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32, _VP, _VP, _VP], ctypes.c_int),
}

_lib = None


class PfError(RuntimeError):
    pass


def load(path=None):
    """Load libpfhip.so (once). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PfError(
            f"{p} is missing: build the HIP extension first "
            "(python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(p)
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def check(status):
    if status != PF_OK:
        msg = _lib.pf_last_error().decode() if _lib is not None else ""
        raise PfError(f"powerfoam_amd call failed with status {status}: {msg}")


def ptr(t):
    """Raw address of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def tiles_hw(height, width):
    return 1 + (height - 1) // TILE_WIDTH, 1 + (width - 1) // TILE_WIDTH
