"""LumenEngine of the Python SDK (python/lumen/client.py:18-219) on liblumen.

RasterizerConfig and RenderResult keep their fields.  LumenEngine.render
calls lumen_rasterize (the 22-function C ABI of lumen.h, which renders with
the GPU splat pipeline) on a fresh framebuffer of config.image_width x
image_height, so RenderResult holds what that call leaves there: 8-bit
colour trunc(255 min(1, C)), alpha = 1 - T and depth = D / alpha (1000 where
nothing was drawn), with the reference's stats keys.  render_batch renders
many cameras of one GPU-resident scene in one lumen_ex_render_batch frame
(CameraTrajectory.orbit feeds it) and returns, per camera, exactly what
render returns.  LumenEngine(legacy=True) selects lumen_rasterize's legacy
mode instead: the reference C library's own isotropic, DC-only renderer, bit
for bit, whose model is the one the reference SDK's pure-Python render loop
follows (python/lumen/client.py:67-219, there in float64).  There is no CPU
fallback: a missing liblumen.so raises.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from .. import _lib
from .camera import Camera
from .cloud import GaussianCloud


@dataclass
class RasterizerConfig:
    image_width: int = 400
    image_height: int = 300
    tile_size: int = 16
    near_plane: float = 0.1
    far_plane: float = 50.0
    early_termination_alpha: float = 0.99
    gaussian_radius_sigma: float = 3.0

    def c_struct(self) -> _lib.LumenConfig:
        c = _lib.LumenConfig()
        c.image_width, c.image_height, c.tile_size = (self.image_width, self.image_height,
                                                      self.tile_size)
        c.near_plane, c.far_plane = self.near_plane, self.far_plane
        c.early_termination_alpha = self.early_termination_alpha
        c.gaussian_radius_sigma = self.gaussian_radius_sigma
        return c

    def gpu_tile(self) -> int:
        """The tile edge lumen_rasterize uses (16 unless 8..16)."""
        ts = self.tile_size
        return ts if 64 <= ts * ts <= 256 else 16


@dataclass
class RenderResult:
    width: int
    height: int
    rgb_bytes: bytes
    depth_values: List[float]
    alpha_values: List[float]
    stats: Dict[str, Any]


def _camera_struct(cam: Camera) -> _lib.LumenCamera:
    c = _lib.LumenCamera()
    c.position = _lib.LumenVec3(*cam.position)
    c.target = _lib.LumenVec3(*cam.target)
    c.up = _lib.LumenVec3(*cam.up)
    c.fov_x_degrees, c.fov_y_degrees = cam.fov_x_deg, cam.fov_y_deg
    c.near_plane, c.far_plane = cam.near_plane, cam.far_plane
    return c


def _stats(total, visible, total_tiles, touched, ms) -> Dict[str, Any]:
    return {"total_splats": int(total), "visible_splats": int(visible),
            "culled_splats": int(total) - int(visible), "total_tiles": int(total_tiles),
            "touched_tiles": int(touched), "elapsed_ms": float(ms),
            "fps": (1000.0 / ms) if ms > 0 else 1000.0}


class LumenEngine:
    def __init__(self, config: Optional[RasterizerConfig] = None, legacy: bool = False):
        self.config = config if config is not None else RasterizerConfig()
        self.legacy = bool(legacy)
        self._c_lib = _lib.lib()
        self._rasterizer = None
        self._batch = None

    def close(self) -> None:
        if self._rasterizer:
            self._c_lib.lumen_rasterizer_destroy(self._rasterizer)
            self._rasterizer = None
        if self._batch is not None:
            self._batch.close()
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _cloud_handle(self, cloud: GaussianCloud):
        L = self._c_lib
        n = len(cloud)
        h = L.lumen_cloud_create(max(n, 1))
        if not h:
            raise MemoryError("lumen_cloud_create failed")
        recs = (_lib.LumenGaussian * n).from_buffer_copy(cloud.aos().tobytes()) if n else []
        for i in range(n):
            st = L.lumen_cloud_add_gaussian(h, ctypes.byref(recs[i]))
            if st != 0:
                L.lumen_cloud_destroy(h)
                raise RuntimeError(f"lumen_cloud_add_gaussian returned {st}")
        return h

    def render(self, cloud: GaussianCloud, camera: Camera) -> RenderResult:
        """One frame through lumen_rasterize on an empty framebuffer."""
        t0 = time.perf_counter()
        L = self._c_lib
        w, h = self.config.image_width, self.config.image_height
        cfg = self.config.c_struct()
        if not self._rasterizer:
            self._rasterizer = L.lumen_rasterizer_create(ctypes.byref(cfg))
            if not self._rasterizer:
                raise MemoryError("lumen_rasterizer_create failed")
            mode = _lib.LUMEN_EX_RASTER_LEGACY if self.legacy else _lib.LUMEN_EX_RASTER_PIPELINE
            _check("lumen_ex_rasterizer_set_mode", L.lumen_ex_rasterizer_set_mode(self._rasterizer, mode))
        else:
            _check("lumen_rasterizer_set_config",
                   L.lumen_rasterizer_set_config(self._rasterizer, ctypes.byref(cfg)))
        handle = self._cloud_handle(cloud)
        fb = L.lumen_framebuffer_create(w, h)
        try:
            if not fb:
                raise MemoryError("lumen_framebuffer_create failed")
            cam = _camera_struct(camera)
            st = _lib.LumenStats()
            _check("lumen_rasterize", L.lumen_rasterize(self._rasterizer, handle, ctypes.byref(cam),
                                                        fb, ctypes.byref(st)))
            P = w * h
            rgb = ctypes.string_at(fb.contents.rgb_data, 3 * P)
            depth = np.ctypeslib.as_array(fb.contents.depth_data, shape=(P,)).tolist()
            alpha = np.ctypeslib.as_array(fb.contents.alpha_data, shape=(P,)).tolist()
            visible, total_tiles, touched = st.visible_splats, st.total_tiles, st.touched_tiles
        finally:
            if fb:
                L.lumen_framebuffer_destroy(fb)
            L.lumen_cloud_destroy(handle)
        ms = (time.perf_counter() - t0) * 1000.0
        return RenderResult(w, h, rgb, depth, alpha, _stats(len(cloud), visible, total_tiles,
                                                            touched, ms))

    def render_batch(self, cloud: GaussianCloud, cameras: Sequence[Camera]) -> List[RenderResult]:
        """render() for every camera, as one multi-view GPU frame."""
        from ..batch import MultiViewRenderer
        from ..raster import RenderConfig
        t0 = time.perf_counter()
        if not cameras:
            return []
        if self.legacy:  # the C library's renderer is one camera per call
            return [self.render(cloud, c) for c in cameras]
        w, h = self.config.image_width, self.config.image_height
        if self._batch is None:
            self._batch = MultiViewRenderer(0)
        R = self._batch
        scene = cloud.to_scene().to_device("cuda:0")
        pcs = [c.to_pinhole(w, h, self.config.near_plane, self.config.far_plane) for c in cameras]
        ts = self.config.gpu_tile()
        cfg = RenderConfig(tile_width=ts, tile_height=ts)
        out = R.render(scene, pcs, cfg, save_pixels=True, want_stats=True)
        R.synchronize()
        P = w * h
        n = len(cameras)
        rgb = out.rgb.view(n, P, 3).cpu().numpy()
        dep = out.depth.view(n, P).cpu().numpy()
        tfin = R.rasterizer.frame_stage("t_final", n * P, np.float32).reshape(n, P)
        tx, ty = cfg.tiles(w, h)
        ranges = R.rasterizer.frame_stage("ranges", 2 * n * tx * ty, np.int32).reshape(n, tx * ty, 2)
        ms = (time.perf_counter() - t0) * 1000.0 / n
        results = []
        for k in range(n):
            a = (np.float32(1.0) - tfin[k]).astype(np.float32)
            hit = a > 0.0
            c8 = (np.minimum(rgb[k], np.float32(1.0)) * np.float32(255.0)).astype(np.uint8)
            c8[~hit] = 0
            z = np.full(P, 1000.0, np.float32)
            with np.errstate(divide="ignore", invalid="ignore"):
                zz = dep[k] / np.where(hit, a, np.float32(1.0))
            near = hit & (zz < np.float32(1000.0))
            z[near] = zz[near]
            alpha = np.where(hit, a, np.float32(0.0)).astype(np.float32)
            r = ranges[k]
            touched = int(np.count_nonzero((r[:, 0] != -1) & (r[:, 0] < r[:, 1])))
            results.append(RenderResult(w, h, c8.tobytes(), z.tolist(), alpha.tolist(),
                                        _stats(len(cloud), out.stats[k].visible_splats,
                                               tx * ty, touched, ms)))
        return results


def _check(fn: str, status: int) -> None:
    if status != 0:
        raise RuntimeError(f"{fn} returned {status}")
