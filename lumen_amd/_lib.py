"""ctypes binding of liblumen.so (include/lumen.h + include/lumen_ex.h).

The library is loaded from this package directory only (built in-tree by
lumen_amd.build).  Loading fails loudly when it is missing: there is no CPU
fallback anywhere on the product path.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; LUMEN_LIBRARY names another build of the same sources
LIB_PATH = os.environ.get("LUMEN_LIBRARY") or os.path.join(_HERE, "liblumen.so")

c_float_p = ctypes.POINTER(ctypes.c_float)


# ---- lumen.h types (c_api/include/lumen.h:39-102) -------------------------
class LumenConfig(ctypes.Structure):
    _fields_ = [("image_width", ctypes.c_uint32), ("image_height", ctypes.c_uint32),
                ("tile_size", ctypes.c_uint32), ("near_plane", ctypes.c_float),
                ("far_plane", ctypes.c_float), ("early_termination_alpha", ctypes.c_float),
                ("gaussian_radius_sigma", ctypes.c_float)]


class LumenVec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class LumenQuat(ctypes.Structure):
    _fields_ = [("w", ctypes.c_float), ("x", ctypes.c_float), ("y", ctypes.c_float),
                ("z", ctypes.c_float)]


class LumenCamera(ctypes.Structure):
    _fields_ = [("position", LumenVec3), ("target", LumenVec3), ("up", LumenVec3),
                ("fov_x_degrees", ctypes.c_float), ("fov_y_degrees", ctypes.c_float),
                ("near_plane", ctypes.c_float), ("far_plane", ctypes.c_float)]


class LumenGaussian(ctypes.Structure):
    _fields_ = [("position", LumenVec3), ("scale", LumenVec3), ("rotation", LumenQuat),
                ("opacity", ctypes.c_float), ("sh_coeffs", ctypes.c_float * 48)]


class LumenFrameBuffer(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("rgb_data", ctypes.POINTER(ctypes.c_uint8)), ("depth_data", c_float_p),
                ("alpha_data", c_float_p)]


class LumenStats(ctypes.Structure):
    _fields_ = [("total_splats", ctypes.c_size_t), ("visible_splats", ctypes.c_size_t),
                ("culled_splats", ctypes.c_size_t), ("total_tiles", ctypes.c_size_t),
                ("touched_tiles", ctypes.c_size_t), ("rasterize_time_ms", ctypes.c_float),
                ("frames_per_second", ctypes.c_float)]


# ---- lumen_ex.h types ------------------------------------------------------
class LumenExScene(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64)] + [
        (n, ctypes.c_void_p) for n in ("pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z",
                                       "rot_w", "rot_x", "rot_y", "rot_z", "opacity", "sh")]


class LumenExCamera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("focal_x", ctypes.c_float), ("focal_y", ctypes.c_float),
                ("principal_x", ctypes.c_float), ("principal_y", ctypes.c_float),
                ("near_plane", ctypes.c_float), ("far_plane", ctypes.c_float),
                ("position", ctypes.c_float * 3), ("rotation", ctypes.c_float * 4)]


class LumenExRenderConfig(ctypes.Structure):
    _fields_ = [("tile_width", ctypes.c_uint32), ("tile_height", ctypes.c_uint32),
                ("lowpass_filter", ctypes.c_float), ("max_sh_degree", ctypes.c_int32),
                ("white_background", ctypes.c_int32), ("compute_depth", ctypes.c_int32),
                ("flags", ctypes.c_uint32), ("batch_depth", ctypes.c_uint32)]


class LumenExRenderStats(ctypes.Structure):
    _fields_ = [("visible_splats", ctypes.c_uint64), ("total_splat_instances", ctypes.c_uint64),
                ("tiles_x", ctypes.c_uint64), ("tiles_y", ctypes.c_uint64),
                ("sort_passes", ctypes.c_uint32), ("key_bytes", ctypes.c_uint32),
                ("emit_passes", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class LumenExViewStats(ctypes.Structure):
    _fields_ = [("vis_base", ctypes.c_uint32), ("visible", ctypes.c_uint32),
                ("key_base", ctypes.c_uint32), ("instances", ctypes.c_uint32)]


class LumenExGradients(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z", "rot_w", "rot_x", "rot_y",
        "rot_z", "opacity", "sh", "viewspace_grad_2d", "observation_count")]


LUMEN_EX_KEEP_STAGES = 1
LUMEN_EX_SAVE_PIXELS = 2
LUMEN_EX_GRAPH = 4
LUMEN_EX_GRAD_ASSIGN = 8
LUMEN_EX_RESERVED = 16
LUMEN_EX_RASTER_PIPELINE = 0
LUMEN_EX_RASTER_LEGACY = 1
LUMEN_EX_ERROR_CAPACITY = -101

STAGES = {"vis_to_orig": 0, "records": 1, "radius": 2, "keys_unsorted": 3, "vals_unsorted": 4,
          "keys": 5, "vals": 6, "ranges": 7, "t_final": 8, "n_eval": 9, "grad_vis": 10}

LEGACY_SYMBOLS = [
    "lumen_get_version", "lumen_get_version_major", "lumen_get_version_minor",
    "lumen_get_version_patch", "lumen_default_config", "lumen_cloud_create", "lumen_cloud_destroy",
    "lumen_cloud_reserve", "lumen_cloud_get_count", "lumen_cloud_get_capacity",
    "lumen_cloud_add_gaussian", "lumen_cloud_get_gaussian", "lumen_cloud_clear",
    "lumen_create_synthetic_cube", "lumen_create_synthetic_sphere", "lumen_framebuffer_create",
    "lumen_framebuffer_destroy", "lumen_framebuffer_clear", "lumen_rasterizer_create",
    "lumen_rasterizer_destroy", "lumen_rasterizer_set_config", "lumen_rasterize",
]

_SIGS = {
    "lumen_get_version": (ctypes.c_char_p, []),
    "lumen_get_version_major": (ctypes.c_int, []),
    "lumen_get_version_minor": (ctypes.c_int, []),
    "lumen_get_version_patch": (ctypes.c_int, []),
    "lumen_default_config": (LumenConfig, []),
    "lumen_cloud_create": (ctypes.c_void_p, [ctypes.c_size_t]),
    "lumen_cloud_destroy": (None, [ctypes.c_void_p]),
    "lumen_cloud_reserve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t]),
    "lumen_cloud_get_count": (ctypes.c_size_t, [ctypes.c_void_p]),
    "lumen_cloud_get_capacity": (ctypes.c_size_t, [ctypes.c_void_p]),
    "lumen_cloud_add_gaussian": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenGaussian)]),
    "lumen_cloud_get_gaussian": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t,
                                                ctypes.POINTER(LumenGaussian)]),
    "lumen_cloud_clear": (None, [ctypes.c_void_p]),
    "lumen_create_synthetic_cube": (ctypes.c_void_p, [ctypes.c_size_t, ctypes.c_float]),
    "lumen_create_synthetic_sphere": (ctypes.c_void_p, [ctypes.c_size_t, ctypes.c_float]),
    "lumen_framebuffer_create": (ctypes.POINTER(LumenFrameBuffer), [ctypes.c_uint32, ctypes.c_uint32]),
    "lumen_framebuffer_destroy": (None, [ctypes.POINTER(LumenFrameBuffer)]),
    "lumen_framebuffer_clear": (None, [ctypes.POINTER(LumenFrameBuffer), ctypes.c_uint8,
                                       ctypes.c_uint8, ctypes.c_uint8]),
    "lumen_rasterizer_create": (ctypes.c_void_p, [ctypes.POINTER(LumenConfig)]),
    "lumen_rasterizer_destroy": (None, [ctypes.c_void_p]),
    "lumen_rasterizer_set_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenConfig)]),
    "lumen_rasterize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(LumenCamera),
                                       ctypes.POINTER(LumenFrameBuffer), ctypes.POINTER(LumenStats)]),
    "lumen_ex_rasterizer_set_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "lumen_ex_rasterizer_get_mode": (ctypes.c_int, [ctypes.c_void_p]),
    "lumen_ex_abi_version": (ctypes.c_int, []),
    "lumen_ex_last_error": (ctypes.c_char_p, []),
    "lumen_ex_context_create": (ctypes.c_void_p, [ctypes.c_int]),
    "lumen_ex_context_destroy": (None, [ctypes.c_void_p]),
    "lumen_ex_render": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenExScene),
                                       ctypes.POINTER(LumenExCamera),
                                       ctypes.POINTER(LumenExRenderConfig), ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.POINTER(LumenExRenderStats),
                                       ctypes.c_void_p]),
    "lumen_ex_radix_sort": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_uint64, ctypes.c_void_p]),
    "lumen_ex_build_ranges": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    "lumen_ex_copy_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_size_t]),
    "lumen_ex_render_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenExScene),
                                                ctypes.POINTER(LumenExCamera),
                                                ctypes.POINTER(LumenExRenderConfig),
                                                ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.POINTER(LumenExGradients),
                                                ctypes.c_void_p]),
    "lumen_ex_render_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenExScene),
                                             ctypes.POINTER(LumenExCamera), ctypes.c_uint32,
                                             ctypes.POINTER(LumenExRenderConfig), ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.POINTER(LumenExViewStats),
                                             ctypes.c_void_p]),
    "lumen_ex_frame_status": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenExViewStats),
                                             ctypes.c_uint32, ctypes.c_void_p]),
    "lumen_ex_render_batch_backward": (ctypes.c_int, [ctypes.c_void_p,
                                                      ctypes.POINTER(LumenExScene),
                                                      ctypes.POINTER(LumenExCamera),
                                                      ctypes.c_uint32,
                                                      ctypes.POINTER(LumenExRenderConfig),
                                                      ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.POINTER(LumenExGradients),
                                                      ctypes.c_void_p]),
    "lumen_ex_photometric_loss": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                                 ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p]),
    "lumen_ex_photometric_loss_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                         ctypes.c_uint32, ctypes.c_uint32,
                                                         ctypes.c_float, ctypes.c_void_p,
                                                         ctypes.c_void_p, ctypes.c_void_p]),
    "lumen_ex_workspace_size": (ctypes.c_int, [ctypes.POINTER(LumenExCamera), ctypes.c_uint32,
                                               ctypes.POINTER(LumenExRenderConfig),
                                               ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_size_t)]),
    "lumen_ex_context_reserve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenExCamera),
                                                ctypes.c_uint32,
                                                ctypes.POINTER(LumenExRenderConfig),
                                                ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]),
    "lumen_ex_context_workspace": (ctypes.c_int, [ctypes.c_void_p,
                                                  ctypes.POINTER(ctypes.c_size_t)]),
    "lumen_ex_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "lumen_ex_stage_times": (ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int]),
}

class LumenExPlanes(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in (
        "pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z", "rot_w", "rot_x", "rot_y",
        "rot_z", "opacity", "sh")]


class LumenExTrainConfig(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64)] + [(k, ctypes.c_float) for k in (
        "position_lr_init", "position_lr_final", "feature_lr", "opacity_lr", "scaling_lr",
        "rotation_lr", "beta1", "beta2", "epsilon", "densify_grad_threshold",
        "prune_opacity_threshold")]


class LumenExDensifyStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("cloned", "split", "pruned", "count")]


LUMEN_EX_ADAM_ALL_PARAMS = 1

_SIGS.update({
    "lumen_ex_adam_step": (ctypes.c_int, [ctypes.POINTER(LumenExPlanes),
                                          ctypes.POINTER(LumenExGradients),
                                          ctypes.POINTER(LumenExPlanes),
                                          ctypes.POINTER(LumenExPlanes), ctypes.c_uint64,
                                          ctypes.POINTER(LumenExTrainConfig), ctypes.c_int64,
                                          ctypes.c_uint32, ctypes.c_void_p]),
    "lumen_ex_reset_opacities": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_float,
                                                ctypes.c_void_p]),
    "lumen_ex_densify_and_prune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LumenExScene),
                                                  ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.POINTER(LumenExTrainConfig),
                                                  ctypes.c_float, ctypes.POINTER(LumenExPlanes),
                                                  ctypes.c_uint64, ctypes.c_void_p,
                                                  ctypes.c_void_p,
                                                  ctypes.POINTER(LumenExDensifyStats),
                                                  ctypes.c_void_p]),
})

TIMED_STAGES = ("count", "emit", "sort", "ranges", "composite", "bwd_composite", "bwd_preprocess")

_lib = None
_lock = threading.Lock()


class LumenError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} returned {status}: {msg}")
        self.status = status


def lib() -> ctypes.CDLL:
    """The loaded liblumen.so with argument types declared."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m lumen_amd.build` "
                    "(there is no fallback path)")
            # PyTorch ships its own HIP runtime under the same soname
            # (libamdhip64.so.7).  Importing it first makes liblumen.so bind
            # to that runtime, so that the process holds one HIP runtime
            # whichever of the two a caller uses first; loaded the other way
            # round, torch would find the system runtime already in place
            # and fail to initialize its devices.
            import torch  # noqa: F401
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
        return _lib


def check(fn: str, status: int) -> None:
    if status != 0:
        msg = lib().lumen_ex_last_error()
        raise LumenError(fn, status, msg.decode() if msg else "")
