"""Differentiable rendering, gradient bookkeeping and the photometric loss.

- GaussianGradients mirrors src/optim/gradients.mojo:10-106 (same fields and
  methods) on device tensors: gradient planes in the GaussianSoA order, SH
  gradients, and the view-space positional gradient accumulator that adaptive
  density control reads (src/optim/density_control.mojo:67).
- RasterizeSplats is a torch.autograd.Function whose forward is lumen_ex_render
  and whose backward is lumen_ex_render_backward (both in liblumen.so).
- l1_loss / ssim_window / combined_loss / psnr restate src/optim/loss.mojo
  with torch ops (used on device, differentiable).
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from .camera import PinholeCamera
from .raster import Rasterizer, RenderConfig, default_rasterizer
from .scene import DeviceScene, PLANE_NAMES


class GaussianGradients:
    """src/optim/gradients.mojo:10-106 on the GPU."""

    def __init__(self, count: int, device="cuda"):
        self.planes = torch.zeros((11, count), dtype=torch.float32, device=device)
        self.sh = torch.zeros((count, 48), dtype=torch.float32, device=device)
        self.viewspace_grad_2d = torch.zeros(count, dtype=torch.float32, device=device)
        self.observation_count = torch.zeros(count, dtype=torch.int32, device=device)

    @property
    def count(self) -> int:
        return int(self.planes.shape[1])

    def __getattr__(self, name):
        # grad_pos_x ... grad_opacity as views of the planes
        if name.startswith("grad_") and name[5:] in PLANE_NAMES:
            return self.planes[PLANE_NAMES.index(name[5:])]
        if name == "grad_sh":
            return self.sh
        raise AttributeError(name)

    def zero(self) -> None:
        """Parameter gradients only, as the reference's zero() (gradients.mojo:75-94)."""
        self.planes.zero_()
        self.sh.zero_()

    def accumulate_viewspace_grad(self, index: int, grad_2d_norm: float) -> None:
        if 0 <= index < self.count:
            self.viewspace_grad_2d[index] += grad_2d_norm
            self.observation_count[index] += 1

    def get_average_viewspace_grad(self, index: int) -> float:
        if 0 <= index < self.count:
            c = int(self.observation_count[index])
            if c > 0:
                return float(self.viewspace_grad_2d[index] / c)
        return 0.0

    def average_viewspace_grad(self) -> torch.Tensor:
        c = self.observation_count.to(torch.float32)
        return torch.where(c > 0, self.viewspace_grad_2d / c.clamp_min(1.0), torch.zeros_like(c))

    def ex_struct(self):
        from ._lib import LumenExGradients
        g = LumenExGradients()
        n = self.count
        base = self.planes.data_ptr()
        for k, name in enumerate(PLANE_NAMES):
            setattr(g, name, base + 4 * k * n)
        g.sh = self.sh.data_ptr()
        g.viewspace_grad_2d = self.viewspace_grad_2d.data_ptr()
        g.observation_count = self.observation_count.data_ptr()
        return g


class RasterizeSplats(torch.autograd.Function):
    """(planes (11, N), sh (N, 48)) -> (rgb (H, W, 3), depth (H, W))."""

    @staticmethod
    def forward(ctx, planes, sh, camera: PinholeCamera, config: RenderConfig,
                rasterizer: Rasterizer, viewspace: Optional[GaussianGradients]):
        scene = DeviceScene(planes.detach(), sh.detach())
        out = rasterizer.render(scene, camera, config, save_pixels=True)
        ctx.scene = scene
        ctx.camera = camera
        ctx.config = config
        ctx.rasterizer = rasterizer
        ctx.viewspace = viewspace
        ctx.frame = rasterizer.frame_id
        return out.rgb, out.depth

    @staticmethod
    def backward(ctx, grad_rgb, grad_depth):
        R: Rasterizer = ctx.rasterizer
        if R.frame_id != ctx.frame:
            # another render used this context since the forward: redo it
            R.render(ctx.scene, ctx.camera, ctx.config, save_pixels=True)
        grads = GaussianGradients(ctx.scene.count, ctx.scene.device)
        if ctx.viewspace is not None:
            # densification statistics accumulate across views in the caller's object
            grads.viewspace_grad_2d = ctx.viewspace.viewspace_grad_2d
            grads.observation_count = ctx.viewspace.observation_count
        g_rgb = grad_rgb.contiguous() if grad_rgb is not None else \
            torch.zeros((ctx.camera.config.height, ctx.camera.config.width, 3),
                        device=ctx.scene.device)
        g_d = grad_depth.contiguous() if grad_depth is not None else None
        R.render_backward(ctx.scene, ctx.camera, ctx.config, g_rgb, g_d, grads)
        return grads.planes, grads.sh, None, None, None, None


def rasterize(planes: torch.Tensor, sh: torch.Tensor, camera: PinholeCamera,
              config: Optional[RenderConfig] = None, rasterizer: Optional[Rasterizer] = None,
              viewspace: Optional[GaussianGradients] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """Differentiable rasterize_scene.  If `viewspace` is given, its
    view-space gradient accumulator (and parameter gradients) also receive
    this view's contribution during backward."""
    return RasterizeSplats.apply(planes, sh, camera, config or RenderConfig(),
                                 rasterizer or default_rasterizer(), viewspace)


# ---- losses: src/optim/loss.mojo ------------------------------------------

SSIM_C1 = 0.0001
SSIM_C2 = 0.0009


def l1_loss(pred: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Mean |pred - target| whose gradient is +1/N where pred >= target and
    -1/N elsewhere (compute_l1_loss, loss.mojo:11-35)."""
    diff = pred - target
    return torch.where(diff >= 0, diff, -diff).mean()


def ssim_window(pred: torch.Tensor, target: torch.Tensor, win: int = 8) -> torch.Tensor:
    """Mean SSIM over non-overlapping win x win windows with the three colour
    channels pooled (compute_ssim_window, loss.mojo:65-133); partial windows
    at the right/bottom edges are not counted.  pred/target: (H, W, 3)."""
    H, W = pred.shape[0], pred.shape[1]
    hy, wx = H // win, W // win
    if hy == 0 or wx == 0:
        return torch.ones((), device=pred.device, dtype=pred.dtype)

    def blocks(x):
        x = x[: hy * win, : wx * win]
        return x.reshape(hy, win, wx, win, 3).permute(0, 2, 1, 3, 4).reshape(hy, wx, win * win * 3)

    x, y = blocks(pred), blocks(target)
    mx, my = x.mean(-1, keepdim=True), y.mean(-1, keepdim=True)
    dx, dy = x - mx, y - my
    vx, vy = (dx * dx).mean(-1), (dy * dy).mean(-1)
    cxy = (dx * dy).mean(-1)
    mx, my = mx[..., 0], my[..., 0]
    num = (2.0 * mx * my + SSIM_C1) * (2.0 * cxy + SSIM_C2)
    den = (mx * mx + my * my + SSIM_C1) * (vx + vy + SSIM_C2)
    return (num / den).mean()


def combined_loss(pred: torch.Tensor, target: torch.Tensor, lambda_dssim: float = 0.2) -> torch.Tensor:
    """(1 - lambda) * L1 + lambda * (1 - SSIM) (compute_combined_loss, loss.mojo:136-144)."""
    return (1.0 - lambda_dssim) * l1_loss(pred, target) + lambda_dssim * (1.0 - ssim_window(pred, target))


def psnr(pred: torch.Tensor, target: torch.Tensor) -> float:
    mse = float(((pred - target) ** 2).mean())
    if mse < 1e-10:
        return 99.99
    import math
    return -4.3429448 * math.log(mse)
