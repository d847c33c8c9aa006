"""Many views of one scene per frame: lumen_ex_render_batch on MI355X.

One launch per stage covers every view (the per-Gaussian kernels loop over
the views with the Gaussian's data in registers; the sort orders all views'
keys at once; the compositor's grid is every view's tiles), and with
graph=True the whole frame is captured once as a hipGraph and replayed.
Each view's image equals a single-view rasterize_scene of that camera.

The autotuner picks the tile shape and the compositor's LDS batch depth for
a (scene, cameras) workload by timing the candidates on the device.
"""
from __future__ import annotations

import ctypes
import itertools
import time
from dataclasses import dataclass, replace
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import _lib
from .camera import PinholeCamera
from .raster import Rasterizer, RenderConfig, camera_struct
from .scene import DeviceScene


@dataclass
class ViewStats:
    vis_base: int
    visible_splats: int
    key_base: int
    total_splat_instances: int


class BatchOutput:
    """Outputs of one frame: flat rgb/depth planes and per-view views of them."""

    def __init__(self, cameras: Sequence[PinholeCamera], rgb: torch.Tensor, depth: torch.Tensor):
        self.cameras = list(cameras)
        self.rgb = rgb
        self.depth = depth
        self.views_rgb: List[torch.Tensor] = []
        self.views_depth: List[torch.Tensor] = []
        base = 0
        for cam in self.cameras:
            W, H = cam.config.width, cam.config.height
            self.views_rgb.append(rgb[3 * base: 3 * (base + W * H)].view(H, W, 3))
            self.views_depth.append(depth[base: base + W * H].view(H, W))
            base += W * H
        self.stats: Optional[List[ViewStats]] = None


def frame_pixels(cameras: Sequence[PinholeCamera]) -> int:
    return sum(c.config.width * c.config.height for c in cameras)


class MultiViewRenderer:
    """Owns a lumen_ex context, a non-default stream (needed for graph capture)
    and the frame's output planes (reused, so graph replays see the same
    pointers)."""

    def __init__(self, device: int = 0, rasterizer: Optional[Rasterizer] = None):
        self.rasterizer = rasterizer or Rasterizer(device)
        self.device = device
        self.stream = torch.cuda.Stream(device=f"cuda:{device}")
        self._outs: Dict[Tuple, Tuple[torch.Tensor, torch.Tensor]] = {}
        self.frame_id = 0

    def close(self) -> None:
        """Releases the context's device buffers and the output planes."""
        self.stream.synchronize()
        self._outs.clear()
        self.rasterizer.close()

    def _new_outputs(self, cameras) -> Tuple[torch.Tensor, torch.Tensor]:
        P = frame_pixels(cameras)
        with torch.cuda.stream(self.stream):
            return (torch.empty(3 * P, dtype=torch.float32, device=f"cuda:{self.device}"),
                    torch.empty(P, dtype=torch.float32, device=f"cuda:{self.device}"))

    def _outputs(self, cameras) -> Tuple[torch.Tensor, torch.Tensor]:
        """The graph frame's output planes: graph replays need fixed pointers,
        so they are kept (for the latest frame shape only) and overwritten by
        the next frame."""
        key = tuple((c.config.width, c.config.height) for c in cameras)
        if key not in self._outs:
            self._outs.clear()
            self._outs[key] = self._new_outputs(cameras)
        return self._outs[key]

    def render(self, scene: DeviceScene, cameras: Sequence[PinholeCamera],
               config: Optional[RenderConfig] = None, graph: bool = False,
               save_pixels: bool = False, want_stats: bool = False,
               out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
               keep_stages: bool = False, reserved: bool = False) -> BatchOutput:
        """Enqueue one frame on self.stream (the caller's current stream is
        made to wait for the scene first).  want_stats synchronizes.
        reserved=True enqueues it with the capacities the context holds and
        returns without any host round trip (LUMEN_EX_RESERVED); status()
        then tells whether it fitted."""
        config = config or RenderConfig()
        n = len(cameras)
        cams = (_lib.LumenExCamera * n)(*[camera_struct(c) for c in cameras])
        flags = (_lib.LUMEN_EX_GRAPH if graph else 0) | \
            (_lib.LUMEN_EX_SAVE_PIXELS if save_pixels else 0) | \
            (_lib.LUMEN_EX_KEEP_STAGES if keep_stages else 0) | \
            (_lib.LUMEN_EX_RESERVED if reserved else 0)
        cfg = config.ex_struct(flags)
        if out is not None:
            rgb, depth = out
        elif graph:
            rgb, depth = self._outputs(cameras)
        else:
            rgb, depth = self._new_outputs(cameras)
        sc = scene.ex_struct()
        self.stream.wait_stream(torch.cuda.current_stream(scene.device))
        stats = (_lib.LumenExViewStats * n)() if want_stats else None
        st = _lib.lib().lumen_ex_render_batch(
            self.rasterizer.handle, ctypes.byref(sc), cams, n, ctypes.byref(cfg), rgb.data_ptr(),
            depth.data_ptr(), stats, self.stream.cuda_stream)
        _lib.check("lumen_ex_render_batch", st)
        self.rasterizer.frame_id += 1
        self.frame_id = self.rasterizer.frame_id
        res = BatchOutput(cameras, rgb, depth)
        if want_stats:
            res.stats = [ViewStats(s.vis_base, s.visible, s.key_base, s.instances) for s in stats]
        return res

    def status(self, n_views: int) -> List[ViewStats]:
        stats = (_lib.LumenExViewStats * n_views)()
        st = _lib.lib().lumen_ex_frame_status(self.rasterizer.handle, stats, n_views,
                                              self.stream.cuda_stream)
        _lib.check("lumen_ex_frame_status", st)
        return [ViewStats(s.vis_base, s.visible, s.key_base, s.instances) for s in stats]

    def synchronize(self) -> None:
        self.stream.synchronize()

    def backward(self, scene: DeviceScene, cameras: Sequence[PinholeCamera], config: RenderConfig,
                 grad_rgb: torch.Tensor, grad_depth: Optional[torch.Tensor], grads,
                 assign: bool = False) -> None:
        """Accumulate (or with assign=True write) the gradient of the last
        frame into `grads` (a GaussianGradients or a LumenExGradients)."""
        n = len(cameras)
        cams = (_lib.LumenExCamera * n)(*[camera_struct(c) for c in cameras])
        cfg = config.ex_struct(_lib.LUMEN_EX_GRAD_ASSIGN if assign else 0)
        sc = scene.ex_struct()
        g = grads if isinstance(grads, _lib.LumenExGradients) else grads.ex_struct()
        self.stream.wait_stream(torch.cuda.current_stream(scene.device))
        st = _lib.lib().lumen_ex_render_batch_backward(
            self.rasterizer.handle, ctypes.byref(sc), cams, n, ctypes.byref(cfg),
            grad_rgb.data_ptr(), grad_depth.data_ptr() if grad_depth is not None else None,
            ctypes.byref(g), self.stream.cuda_stream)
        _lib.check("lumen_ex_render_batch_backward", st)
        torch.cuda.current_stream(scene.device).wait_stream(self.stream)


class RasterizeViews(torch.autograd.Function):
    """(planes, sh) -> flat rgb (3 * sum W*H), flat depth (sum W*H) of every view."""

    @staticmethod
    def forward(ctx, planes, sh, cameras, config, renderer: MultiViewRenderer, graph: bool,
                viewspace, reserved: bool = False):
        # an unused output's gradient arrives as None, not as a zero-filled image
        ctx.set_materialize_grads(False)
        scene = DeviceScene(planes.detach(), sh.detach())
        if graph:
            # replays need fixed output pointers: the renderer's planes are
            # reused (and overwritten) by the next frame
            rgb, depth = renderer._outputs(cameras)
        else:
            P = frame_pixels(cameras)
            rgb = torch.empty(3 * P, dtype=torch.float32, device=planes.device)
            depth = torch.empty(P, dtype=torch.float32, device=planes.device)
        renderer.render(scene, cameras, config, graph=graph, save_pixels=True, out=(rgb, depth),
                        reserved=reserved)
        torch.cuda.current_stream(planes.device).wait_stream(renderer.stream)
        if graph:
            # fresh tensor objects over the renderer's fixed planes, so that the
            # next frame's outputs are other autograd outputs
            rgb, depth = rgb.view(-1), depth.view(-1)
        ctx.scene, ctx.cameras, ctx.config, ctx.renderer = scene, cameras, config, renderer
        ctx.viewspace = viewspace
        ctx.frame = renderer.rasterizer.frame_id
        ctx.graph = graph
        ctx.reserved = reserved
        # outputs go through save_for_backward: kept on ctx directly they would
        # form a cycle (output -> grad_fn -> ctx -> output) that is never freed
        ctx.save_for_backward(rgb, depth)
        return rgb, depth

    @staticmethod
    def backward(ctx, grad_rgb, grad_depth):
        from .training import new_gradients
        R = ctx.renderer
        out_rgb, out_depth = ctx.saved_tensors
        if R.rasterizer.frame_id != ctx.frame:
            R.render(ctx.scene, ctx.cameras, ctx.config, graph=ctx.graph, save_pixels=True,
                     out=(out_rgb, out_depth), reserved=ctx.reserved)
        planes, sh, gs = new_gradients(ctx.scene.count, ctx.scene.device, ctx.viewspace)
        g_rgb = grad_rgb.contiguous() if grad_rgb is not None else torch.zeros_like(out_rgb)
        g_d = grad_depth.contiguous() if grad_depth is not None else None
        R.backward(ctx.scene, ctx.cameras, ctx.config, g_rgb, g_d, gs, assign=True)
        return planes, sh, None, None, None, None, None, None


def rasterize_views(planes, sh, cameras, config=None, renderer=None, graph=False, viewspace=None,
                    reserved=False):
    return RasterizeViews.apply(planes, sh, list(cameras), config or RenderConfig(),
                                renderer or MultiViewRenderer(), graph, viewspace, reserved)


# ---- autotuner ----------------------------------------------------------------

TILE_SHAPES = ((16, 16), (8, 8), (32, 8))
BATCH_DEPTHS = (64, 128, 256)


@dataclass
class TuneResult:
    config: RenderConfig
    ms: float
    table: Dict[str, float]


def autotune(renderer: MultiViewRenderer, scene: DeviceScene, cameras: Sequence[PinholeCamera],
             base: Optional[RenderConfig] = None, graph: bool = True, reps: int = 3,
             shapes=TILE_SHAPES, depths=BATCH_DEPTHS) -> TuneResult:
    """Times every (tile shape, LDS batch depth) on the actual workload and
    returns the fastest.  Images do not depend on either choice."""
    base = base or RenderConfig()
    table = {}
    best = None
    for (tw, th), bd in itertools.product(shapes, depths):
        cfg = replace(base, tile_width=tw, tile_height=th, batch_depth=bd)
        renderer.render(scene, cameras, cfg, graph=graph)   # plan / capture
        renderer.render(scene, cameras, cfg, graph=graph)   # warm
        renderer.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            renderer.render(scene, cameras, cfg, graph=graph)
        renderer.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        table[f"{tw}x{th}/{bd}"] = round(ms, 4)
        if best is None or ms < best[1]:
            best = (cfg, ms)
    return TuneResult(best[0], best[1], table)
