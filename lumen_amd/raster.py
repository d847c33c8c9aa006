"""rasterize_scene on MI355X: the Python face of lumen_ex_render.

Mirrors the reference's interface (src/raster/pipeline.mojo:21-231,
src/core/config.mojo:65-102): RenderConfig with the same field names and
defaults, RenderOutput with rgb / depth buffers and the visible_splats /
total_splat_instances counters, rasterize_scene(cloud, camera, config).
Buffers are torch tensors on the GPU; all computation happens in liblumen.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .camera import PinholeCamera
from .scene import DeviceScene, GaussianScene


@dataclass
class RenderConfig:
    tile_width: int = 16
    tile_height: int = 16
    lowpass_filter: float = 0.3
    max_sh_degree: int = 3
    white_background: bool = False
    compute_depth: bool = True
    batch_depth: int = 0  # compositor LDS batch (64, 128, 256); 0: the library's choice

    def ex_struct(self, flags: int = 0):
        c = _lib.LumenExRenderConfig()
        c.tile_width = self.tile_width
        c.tile_height = self.tile_height
        c.lowpass_filter = self.lowpass_filter
        c.max_sh_degree = self.max_sh_degree
        c.white_background = int(bool(self.white_background))
        c.compute_depth = int(bool(self.compute_depth))
        c.flags = flags
        c.batch_depth = self.batch_depth
        return c

    def tiles(self, width: int, height: int):
        return ((width + self.tile_width - 1) // self.tile_width,
                (height + self.tile_height - 1) // self.tile_height)


def camera_struct(camera: PinholeCamera):
    c = _lib.LumenExCamera()
    cfg = camera.config
    c.width, c.height = cfg.width, cfg.height
    c.focal_x, c.focal_y = cfg.focal_x, cfg.focal_y
    c.principal_x, c.principal_y = cfg.principal_x, cfg.principal_y
    c.near_plane, c.far_plane = cfg.near_plane, cfg.far_plane
    for k in range(3):
        c.position[k] = float(camera.position[k])
    for k in range(4):
        c.rotation[k] = float(camera.rotation[k])
    return c


@dataclass
class RenderOutput:
    width: int
    height: int
    rgb: "object"    # torch.Tensor (H, W, 3) float32
    depth: "object"  # torch.Tensor (H, W) float32
    visible_splats: int
    total_splat_instances: int
    sort_passes: int = 0
    key_bytes: int = 8
    emit_passes: int = 0  # of sort_passes, run while the keys were emitted

    def get_pixel_rgb(self, x: int, y: int):
        if 0 <= x < self.width and 0 <= y < self.height:
            return tuple(float(v) for v in self.rgb[y, x].tolist())
        return (0.0, 0.0, 0.0)


def _stream_handle(device) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream


class Rasterizer:
    """Owns one lumen_ex context (device workspace that grows, never shrinks)."""

    def __init__(self, device: int = 0):
        self._ctx = _lib.lib().lumen_ex_context_create(device)
        if not self._ctx:
            raise RuntimeError("lumen_ex_context_create failed: " +
                               _lib.lib().lumen_ex_last_error().decode())
        self.device = device
        self._last = None
        self.frame_id = 0  # incremented by every render on this context

    def close(self):
        if self._ctx:
            _lib.lib().lumen_ex_context_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._ctx

    def render(self, scene, camera: PinholeCamera, config: Optional[RenderConfig] = None,
               keep_stages: bool = False, save_pixels: bool = False,
               out_rgb=None, out_depth=None) -> RenderOutput:
        import torch
        config = config or RenderConfig()
        if isinstance(scene, GaussianScene):
            scene = scene.to_device(f"cuda:{self.device}")
        dev = scene.device
        W, H = camera.config.width, camera.config.height
        if out_rgb is None:
            out_rgb = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
        if out_depth is None:
            out_depth = torch.empty((H, W), dtype=torch.float32, device=dev)
        assert out_rgb.is_contiguous() and out_rgb.shape == (H, W, 3)
        assert out_depth.is_contiguous() and out_depth.shape == (H, W)
        flags = (_lib.LUMEN_EX_KEEP_STAGES if keep_stages else 0) | \
                (_lib.LUMEN_EX_SAVE_PIXELS if save_pixels else 0)
        sc = scene.ex_struct()
        cam = camera_struct(camera)
        cfg = config.ex_struct(flags)
        stats = _lib.LumenExRenderStats()
        st = _lib.lib().lumen_ex_render(self._ctx, ctypes.byref(sc), ctypes.byref(cam),
                                        ctypes.byref(cfg), out_rgb.data_ptr(),
                                        out_depth.data_ptr(), ctypes.byref(stats),
                                        _stream_handle(dev))
        _lib.check("lumen_ex_render", st)
        self.frame_id += 1
        self._last = (int(stats.visible_splats), int(stats.total_splat_instances),
                      int(stats.tiles_x * stats.tiles_y), W * H)
        return RenderOutput(W, H, out_rgb, out_depth, int(stats.visible_splats),
                            int(stats.total_splat_instances), int(stats.sort_passes),
                            int(stats.key_bytes), int(stats.emit_passes))

    def stage(self, name: str) -> np.ndarray:
        """Host copy of a stage buffer of the last render (see lumen_ex.h)."""
        V, M, T, P = self._last
        shapes = {"grad_vis": (np.float32, (V, 10)),
                  "vis_to_orig": (np.int32, (V,)), "records": (np.float32, (V, 12)),
                  "radius": (np.int32, (V,)), "keys_unsorted": (np.uint64, (M,)),
                  "vals_unsorted": (np.uint32, (M,)), "keys": (np.uint64, (M,)),
                  "vals": (np.uint32, (M,)), "ranges": (np.int32, (T, 2)),
                  "t_final": (np.float32, (P,)), "n_eval": (np.int32, (P,))}
        dt, shape = shapes[name]
        out = np.empty(shape, dt)
        st = _lib.lib().lumen_ex_copy_stage(self._ctx, _lib.STAGES[name],
                                            out.ctypes.data_as(ctypes.c_void_p), out.nbytes)
        _lib.check("lumen_ex_copy_stage", st)
        return out

    def frame_stage(self, name: str, count: int, dtype) -> np.ndarray:
        """Host copy of `count` elements of a stage buffer of the last frame
        (single or multi-view) on this context."""
        out = np.empty(count, dtype)
        st = _lib.lib().lumen_ex_copy_stage(self._ctx, _lib.STAGES[name],
                                            out.ctypes.data_as(ctypes.c_void_p), out.nbytes)
        _lib.check("lumen_ex_copy_stage", st)
        return out

    def render_backward(self, scene, camera: PinholeCamera, config: RenderConfig, grad_rgb,
                        grad_depth, grads, assign: bool = False) -> None:
        """Accumulate the gradient of the last render (made with save_pixels=True
        from the same scene/camera/config) into `grads` (a GaussianGradients or
        a LumenExGradients struct), or write it there with assign=True."""
        assert grad_rgb.is_contiguous() and grad_rgb.dtype.is_floating_point
        sc = scene.ex_struct()
        cam = camera_struct(camera)
        cfg = config.ex_struct(_lib.LUMEN_EX_GRAD_ASSIGN if assign else 0)
        g = grads if isinstance(grads, _lib.LumenExGradients) else grads.ex_struct()
        gd = grad_depth.data_ptr() if grad_depth is not None else None
        st = _lib.lib().lumen_ex_render_backward(self._ctx, ctypes.byref(sc), ctypes.byref(cam),
                                                 ctypes.byref(cfg), grad_rgb.data_ptr(), gd,
                                                 ctypes.byref(g), _stream_handle(scene.device))
        _lib.check("lumen_ex_render_backward", st)

    def reserve(self, n_gaussians: int, cameras, config: Optional[RenderConfig] = None,
                max_instances: int = 0, backward: bool = False, save_pixels: bool = False,
                graph: bool = False) -> None:
        """Allocate the workspace of frames up to these bounds now, so that
        those frames allocate nothing (lumen_ex_context_reserve)."""
        cams, n, cfg = _workspace_args(cameras, config, save_pixels, graph)
        st = _lib.lib().lumen_ex_context_reserve(self._ctx, cams, n, ctypes.byref(cfg),
                                                 int(n_gaussians), int(max_instances),
                                                 int(bool(backward)))
        _lib.check("lumen_ex_context_reserve", st)

    def workspace_bytes(self) -> int:
        """Device bytes the context holds (lumen_ex_context_workspace)."""
        out = ctypes.c_size_t(0)
        _lib.check("lumen_ex_context_workspace",
                   _lib.lib().lumen_ex_context_workspace(self._ctx, ctypes.byref(out)))
        return int(out.value)

    def set_profiling(self, enable: bool) -> None:
        _lib.check("lumen_ex_set_profiling",
                   _lib.lib().lumen_ex_set_profiling(self._ctx, int(bool(enable))))

    def stage_times(self) -> dict:
        """GPU ms per stage of the last render (HIP events on its stream)."""
        n = len(_lib.TIMED_STAGES)
        buf = (ctypes.c_float * n)()
        _lib.check("lumen_ex_stage_times", _lib.lib().lumen_ex_stage_times(self._ctx, buf, n))
        return {k: float(buf[i]) for i, k in enumerate(_lib.TIMED_STAGES)}

    def radix_sort(self, keys, values) -> None:
        """In-place stable sort of int64-stored uint64 keys with int32 values."""
        import torch
        assert keys.dtype == torch.int64 and values.dtype == torch.int32
        assert keys.is_contiguous() and values.is_contiguous() and keys.numel() == values.numel()
        st = _lib.lib().lumen_ex_radix_sort(self._ctx, keys.data_ptr(), values.data_ptr(),
                                            keys.numel(), _stream_handle(keys.device))
        _lib.check("lumen_ex_radix_sort", st)


def _workspace_args(cameras, config, save_pixels, graph):
    if isinstance(cameras, PinholeCamera):
        cameras = [cameras]
    n = len(cameras)
    cams = (_lib.LumenExCamera * n)(*[camera_struct(c) for c in cameras])
    flags = (_lib.LUMEN_EX_SAVE_PIXELS if save_pixels else 0) | (_lib.LUMEN_EX_GRAPH if graph else 0)
    return cams, n, (config or RenderConfig()).ex_struct(flags)


def workspace_size(n_gaussians: int, cameras, config: Optional[RenderConfig] = None,
                   max_instances: int = 0, backward: bool = False, save_pixels: bool = False,
                   graph: bool = False) -> int:
    """Device bytes a context holds for frames of n_gaussians splats over these
    cameras with at most max_instances tile instances (lumen_ex_workspace_size;
    host-only)."""
    cams, n, cfg = _workspace_args(cameras, config, save_pixels, graph)
    out = ctypes.c_size_t(0)
    st = _lib.lib().lumen_ex_workspace_size(cams, n, ctypes.byref(cfg), int(n_gaussians),
                                            int(max_instances), int(bool(backward)),
                                            ctypes.byref(out))
    _lib.check("lumen_ex_workspace_size", st)
    return int(out.value)


def build_ranges(sorted_keys, tiles: int):
    import torch
    ranges = torch.empty((tiles, 2), dtype=torch.int32, device=sorted_keys.device)
    st = _lib.lib().lumen_ex_build_ranges(sorted_keys.data_ptr(), sorted_keys.numel(), tiles,
                                          ranges.data_ptr(), _stream_handle(sorted_keys.device))
    _lib.check("lumen_ex_build_ranges", st)
    return ranges


_default: Optional[Rasterizer] = None


def default_rasterizer() -> Rasterizer:
    global _default
    if _default is None:
        _default = Rasterizer(0)
    return _default


def rasterize_scene(cloud, camera: PinholeCamera, config: Optional[RenderConfig] = None) -> RenderOutput:
    """rasterize_scene(cloud, camera, config) (src/raster/pipeline.mojo:48)."""
    return default_rasterizer().render(cloud, camera, config)
