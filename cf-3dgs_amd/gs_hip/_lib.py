"""ctypes binding of libgs_hip.so (C ABI declared in include/gs_raster.h).

The library is loaded from this directory.  There is no fallback: if the
library is missing, importing the rasteriser fails with instructions to
build it, and every entry point requires tensors on a HIP device.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgs_hip.so")

c_int = ctypes.c_int
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t
c_void_p = ctypes.c_void_p
c_longlong = ctypes.c_longlong
c_double = ctypes.c_double

GS_OK = 0
GS_STATUS_CAPACITY = 1
GS_STATUS_PREFILTERED = 2
GS_RB_COLOR_ONLY = 2  # gs_multi_render_bwd flag: no frame of the set has a depth or alpha cotangent
GS_RB_POSE_ONLY = 4   # gs_multi_render_bwd flag: the frames take the pose-only backward
TILE = 16


class GsFrame(ctypes.Structure):
    _fields_ = [
        ("width", c_int), ("height", c_int), ("tanfovx", c_float), ("tanfovy", c_float),
        ("scale_modifier", c_float), ("sh_degree", c_int), ("prefiltered", c_int),
        ("viewmatrix", c_void_p), ("projmatrix", c_void_p), ("campos", c_void_p), ("bg", c_void_p),
        ("pose", c_void_p), ("binning", c_int), ("deterministic", c_int),
    ]


BINNING_DEPTH_FIRST = 0
BINNING_SORT64 = 1


class GsGaussians(ctypes.Structure):
    _fields_ = [
        ("P", c_int), ("M", c_int), ("means3D", c_void_p), ("shs", c_void_p), ("colors_precomp", c_void_p),
        ("opacities", c_void_p), ("scales", c_void_p), ("rotations", c_void_p), ("cov3D_precomp", c_void_p),
    ]


class GsGrads(ctypes.Structure):
    _fields_ = [
        ("dL_dmeans2D", c_void_p), ("dL_dcolors", c_void_p), ("dL_dopacity", c_void_p),
        ("dL_dmeans3D", c_void_p), ("dL_dcov3D", c_void_p), ("dL_dsh", c_void_p), ("dL_dscales", c_void_p),
        ("dL_drotations", c_void_p), ("dL_dpose", c_void_p),
    ]


class GsAdamFused(ctypes.Structure):
    _fields_ = [
        ("means_m", c_void_p), ("means_v", c_void_p), ("means_lr", c_float),
        ("shs_m", c_void_p), ("shs_v", c_void_p), ("shs_lr", c_float), ("shs_lead_lr", c_float),
        ("shs_lead", c_int), ("beta1", c_float), ("beta2", c_float), ("eps", c_float), ("step", c_int),
    ]


class GsAdamAct(ctypes.Structure):
    _fields_ = [
        ("scale_m", c_void_p), ("scale_v", c_void_p), ("scale_lr", c_float),
        ("rot_m", c_void_p), ("rot_v", c_void_p), ("rot_lr", c_float),
        ("opacity_m", c_void_p), ("opacity_v", c_void_p), ("opacity_lr", c_float),
        ("beta1", c_float), ("beta2", c_float), ("eps", c_float), ("step", c_int),
    ]


class GsGeomLayout(ctypes.Structure):
    _fields_ = [("splat", c_size_t), ("cov3D", c_size_t), ("tiles", c_size_t), ("offsets", c_size_t),
                ("scan_tmp", c_size_t), ("total", c_size_t)]


class GsBinningLayout(ctypes.Structure):
    _fields_ = [("keys_dup", c_size_t), ("vals_dup", c_size_t), ("keys_alt", c_size_t), ("vals_alt", c_size_t),
                ("keys_sorted", c_size_t), ("vals_sorted", c_size_t), ("hist", c_size_t), ("sort_passes", c_int),
                ("key_bytes", c_int), ("id_bits", c_int), ("dkeys_a", c_size_t), ("dvals_a", c_size_t), ("dkeys_b", c_size_t),
                ("dvals_b", c_size_t), ("dscan", c_size_t), ("dblocks", c_size_t), ("drect", c_size_t),
                ("ckpt", c_size_t),
                ("bweight", c_size_t), ("items", c_size_t), ("item_log", c_size_t), ("total", c_size_t)]


class GsImageLayout(ctypes.Structure):
    _fields_ = [("ranges", c_size_t), ("final_T", c_size_t), ("n_contrib", c_size_t), ("accum", c_size_t),
                ("ckpt_base", c_size_t), ("tile_maxc", c_size_t), ("tile_order", c_size_t),
                ("total", c_size_t)]


EXPORTS = {
    "gs_geom_layout": (c_int, [c_int, ctypes.POINTER(GsGeomLayout)]),
    "gs_binning_layout": (c_int, [c_longlong, c_int, c_int, c_int, c_int, ctypes.POINTER(GsBinningLayout)]),
    "gs_image_layout": (c_int, [c_int, c_int, ctypes.POINTER(GsImageLayout)]),
    "gs_backward_scratch_bytes": (c_size_t, [c_int, c_longlong, c_int]),
    "gs_preprocess": (c_int, [ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians), c_void_p, c_void_p, c_size_t,
                              c_void_p, c_void_p]),
    "gs_render": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong, c_void_p, c_size_t,
                          c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gs_render_prepare": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong, c_void_p,
                                  c_size_t, c_void_p, c_size_t, c_void_p]),
    "gs_render_resume": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong, c_void_p,
                                 c_size_t, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gs_backward": (c_int, [ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians), c_void_p, c_void_p, c_void_p,
                            c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(GsGrads),
                            c_void_p, c_size_t, c_void_p]),
    "gs_backward_adam": (c_int, [ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians), c_void_p, c_void_p,
                                 c_void_p, c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 ctypes.POINTER(GsGrads), ctypes.POINTER(GsAdamFused), c_void_p, c_size_t,
                                 c_void_p]),
    "gs_render_binning": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong, c_void_p,
                                  c_size_t, c_void_p, c_size_t, c_void_p]),
    "gs_multi_blend_bytes": (c_size_t, [c_int]),
    "gs_multi_blend_build": (c_int, [c_int, ctypes.POINTER(GsFrame), ctypes.POINTER(c_int), ctypes.POINTER(c_void_p),
                                     ctypes.POINTER(c_longlong), ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                     ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                     c_void_p, c_size_t, ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_blend": (c_int, [c_int, c_void_p, ctypes.c_uint, c_void_p]),
    "gs_render_depth_keys": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong, c_void_p,
                                     c_size_t, c_void_p, c_size_t, c_void_p]),
    "gs_multi_depth_sort_bytes": (c_size_t, [c_int]),
    "gs_multi_depth_sort_build": (c_int, [c_int, ctypes.POINTER(GsFrame), ctypes.POINTER(c_int),
                                          ctypes.POINTER(c_longlong), ctypes.POINTER(c_void_p),
                                          ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p), c_void_p, c_size_t,
                                          ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_depth_sort": (c_int, [c_int, c_void_p, ctypes.c_uint, c_void_p]),
    "gs_render_binning_after_depth_sort": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong,
                                                   c_void_p, c_size_t, c_void_p, c_size_t, c_void_p]),
    "gs_render_binning_duplicate": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong,
                                            c_void_p, c_size_t, c_void_p, c_size_t, c_void_p]),
    "gs_multi_list_sort_bytes": (c_size_t, [c_int]),
    "gs_multi_list_sort_build": (c_int, [c_int, ctypes.POINTER(GsFrame), ctypes.POINTER(c_int),
                                         ctypes.POINTER(c_longlong), ctypes.POINTER(c_void_p),
                                         ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p),
                                         ctypes.POINTER(c_void_p), c_void_p, c_size_t,
                                         ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_list_sort": (c_int, [c_int, c_void_p, ctypes.POINTER(ctypes.c_uint), c_void_p]),
    "gs_render_binning_after_list_sort": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong,
                                                  c_void_p, c_size_t, c_void_p, c_size_t, c_void_p]),
    "gs_render_after_list_sort": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong,
                                          c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p,
                                          c_void_p]),
    "gs_render_after_depth_sort": (c_int, [ctypes.POINTER(GsFrame), c_int, c_void_p, c_void_p, c_longlong,
                                           c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p,
                                           c_void_p]),
    "gs_backward_prepare": (c_int, [ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians), c_void_p, c_void_p,
                                    c_void_p, c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    ctypes.POINTER(GsGrads), c_void_p, c_size_t, c_void_p]),
    "gs_backward_finish": (c_int, [ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians), c_void_p, c_void_p,
                                   c_void_p, c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   ctypes.POINTER(GsGrads), c_void_p, c_size_t, c_void_p]),
    "gs_multi_render_bwd_bytes": (c_size_t, [c_int]),
    "gs_multi_render_bwd_build": (c_int, [c_int, ctypes.POINTER(GsFrame), ctypes.POINTER(c_int),
                                          ctypes.POINTER(c_void_p), ctypes.POINTER(c_longlong),
                                          ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                          ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                          ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_void_p, c_size_t,
                                          ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_render_bwd": (c_int, [c_int, c_void_p, ctypes.c_uint, c_int, c_void_p]),
    "gs_multi_preprocess_keys_bytes": (c_size_t, [c_int]),
    "gs_multi_preprocess_keys_build": (c_int, [c_int, ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians),
                                               ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                               ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p),
                                               ctypes.POINTER(c_longlong), ctypes.POINTER(c_void_p),
                                               ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p),
                                               ctypes.POINTER(c_size_t), c_void_p, c_size_t,
                                               ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_preprocess_keys": (c_int, [c_int, c_void_p, ctypes.POINTER(ctypes.c_uint), c_void_p]),
    "gs_multi_binning_steps_bytes": (c_size_t, [c_int]),
    "gs_multi_binning_steps_build": (c_int, [c_int, ctypes.POINTER(GsFrame), ctypes.POINTER(c_int),
                                             ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                             ctypes.POINTER(c_longlong), ctypes.POINTER(c_void_p),
                                             ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p),
                                             ctypes.POINTER(c_size_t), c_void_p, c_size_t,
                                             ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_binning_duplicate": (c_int, [c_int, c_void_p, ctypes.c_uint, c_void_p]),
    "gs_multi_binning_after_list_sort": (c_int, [c_int, c_void_p, c_void_p]),
    "gs_multi_pose_bwd_bytes": (c_size_t, [c_int]),
    "gs_multi_pose_bwd_build": (c_int, [c_int, ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians),
                                        ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                        ctypes.POINTER(c_void_p), ctypes.POINTER(c_longlong),
                                        ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                        ctypes.POINTER(c_void_p), ctypes.POINTER(GsGrads), ctypes.POINTER(c_void_p),
                                        ctypes.POINTER(c_size_t), c_void_p, c_size_t,
                                        ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_pose_bwd_prepare": (c_int, [c_int, c_void_p, ctypes.POINTER(ctypes.c_uint), c_void_p]),
    "gs_multi_pose_bwd_finish": (c_int, [c_int, c_void_p, ctypes.POINTER(ctypes.c_uint), c_void_p]),
    "gs_backward_is_pose_only": (c_int, [ctypes.POINTER(GsFrame), ctypes.POINTER(GsGaussians),
                                         ctypes.POINTER(GsGrads), c_int]),
    "gs_mark_visible": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gs_sort_tmp_bytes": (c_size_t, [c_longlong]),
    "gs_sort_pairs_u64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_longlong, c_int, c_void_p,
                                  c_size_t, ctypes.POINTER(c_int), c_void_p]),
    "gs_sort_pairs_u32": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_longlong, c_int, c_void_p,
                                  c_size_t, ctypes.POINTER(c_int), c_void_p]),
    "gs_knn_workspace_bytes": (c_size_t, [c_int]),
    "gs_knn3_meandist": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "gs_knn3_meandist_static": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "gs_loss_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "gs_ssim_l1_forward": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, ctypes.POINTER(c_float), c_float,
                                   c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "gs_ssim_l1_backward": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, ctypes.POINTER(c_float), c_float,
                                    c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
    "gs_ssim_l1_forward_backward": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, ctypes.POINTER(c_float),
                                            c_float, c_int, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p,
                                            c_void_p]),
    "gs_multi_loss_bytes": (c_size_t, [c_int]),
    "gs_multi_loss_build": (c_int, [c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), ctypes.POINTER(c_int),
                                    ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_float), c_float,
                                    c_int, ctypes.POINTER(c_void_p), c_void_p, ctypes.POINTER(c_void_p),
                                    ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p), c_void_p, c_size_t,
                                    ctypes.POINTER(ctypes.c_uint)]),
    "gs_multi_loss": (c_int, [c_int, c_void_p, ctypes.c_uint, c_void_p]),
    "gs_se3_retract": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gs_se3_grad": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gs_se3_adam": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float,
                            c_float, c_float, c_void_p, c_void_p]),
    "gs_adam_step": (c_int, [c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                             ctypes.POINTER(c_void_p), ctypes.POINTER(c_longlong), ctypes.POINTER(c_float), c_float,
                             c_float, c_float, c_int, c_void_p]),
    "gs_adam_step_rows": (c_int, [c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                  ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), ctypes.POINTER(c_longlong),
                                  ctypes.POINTER(c_float), ctypes.POINTER(c_float), ctypes.POINTER(c_int),
                                  ctypes.POINTER(c_int), c_float, c_float, c_float, c_int, c_void_p]),
    "gs_activate_forward": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gs_activate_backward": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]),
    "gs_depth_to_points": (c_int, [c_int, c_int, c_void_p, c_float, c_float, c_float, c_float, c_void_p, c_void_p,
                                   c_int, c_longlong, c_longlong, c_void_p, c_void_p]),
    "gs_bounds_workspace_bytes": (c_size_t, []),
    "gs_points_bounds": (c_int, [c_void_p, c_longlong, c_void_p, c_void_p, c_size_t, c_void_p]),
    "gs_voxel_workspace_bytes": (c_size_t, [c_longlong]),
    "gs_voxel_down_sample": (c_int, [c_void_p, c_longlong, c_void_p, c_int, c_longlong, c_longlong,
                                     ctypes.POINTER(c_double), c_double, c_int, c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_size_t, c_void_p]),
    "gs_create_from_pcd": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    "gs_activate_backward_adam": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, ctypes.POINTER(GsAdamAct), c_void_p]),
    "gs_densify_stats": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gs_densify_workspace_bytes": (c_size_t, [c_int]),
    "gs_densify_mark": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                                c_float, c_float, c_int, c_int, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "gs_densify_emit": (c_int, [c_int, c_int, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_void_p,
                                c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "gs_timing_enable": (c_int, [c_int]),
    "gs_timing_reset": (c_int, []),
    "gs_timing_read": (c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_int)]),
    "gs_error_string": (ctypes.c_char_p, [c_int]),
    "gs_abi_version": (c_int, []),
}

_lib = None


class GsError(RuntimeError):
    pass


def load():
    """Load libgs_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python cf-3dgs_amd/build.py` "
                "(hipcc, gfx950). There is no CPU fallback.")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code: int, what: str) -> None:
    if code != GS_OK:
        msg = load().gs_error_string(code).decode()
        raise GsError(f"{what} failed: {msg} (code {code})")


def geom_layout(P: int) -> GsGeomLayout:
    L = GsGeomLayout()
    check(load().gs_geom_layout(P, ctypes.byref(L)), "gs_geom_layout")
    return L


def binning_layout(capacity: int, P: int, W: int, H: int, binning: int = BINNING_DEPTH_FIRST) -> GsBinningLayout:
    L = GsBinningLayout()
    check(load().gs_binning_layout(capacity, P, W, H, binning, ctypes.byref(L)), "gs_binning_layout")
    return L


def image_layout(W: int, H: int) -> GsImageLayout:
    L = GsImageLayout()
    check(load().gs_image_layout(W, H, ctypes.byref(L)), "gs_image_layout")
    return L


STAGES = ("preprocess", "scan", "duplicate", "sort", "ranges", "blend", "render_bwd", "preprocess_bwd", "knn")


def timing_enable(on: bool) -> None:
    check(load().gs_timing_enable(int(bool(on))), "gs_timing_enable")


def timing_reset() -> None:
    check(load().gs_timing_reset(), "gs_timing_reset")


def timing_read(stage: str):
    """(total milliseconds, launches) recorded for `stage` since the last reset."""
    t = ctypes.c_double(0.0)
    n = c_int(0)
    check(load().gs_timing_read(stage.encode(), ctypes.byref(t), ctypes.byref(n)), "gs_timing_read")
    return t.value, n.value
