"""Differentiable rendering, gradient bookkeeping and the photometric loss.

- GaussianGradients mirrors src/optim/gradients.mojo:10-106 (same fields and
  methods) on device tensors: gradient planes in the GaussianSoA order, SH
  gradients, and the view-space positional gradient accumulator that adaptive
  density control reads (src/optim/density_control.mojo:67).
- RasterizeSplats is a torch.autograd.Function whose forward is lumen_ex_render
  and whose backward is lumen_ex_render_backward (both in liblumen.so).
- photometric_loss / l1_loss / ssim_window / combined_loss restate
  src/optim/loss.mojo through lumen_ex_photometric_loss (values and
  gradient in one HIP pass); psnr is a host-side metric.
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

    @classmethod
    def wrap(cls, planes, sh, viewspace_grad_2d, observation_count) -> "GaussianGradients":
        """Gradients over existing tensors (no allocation); planes / sh may be
        None for a holder of the view-space statistics only."""
        g = cls.__new__(cls)
        g.planes, g.sh = planes, sh
        g.viewspace_grad_2d, g.observation_count = viewspace_grad_2d, observation_count
        return g

    @property
    def count(self) -> int:
        return int(self.viewspace_grad_2d.shape[0])

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
        return gradients_struct(self.planes, self.sh, self.viewspace_grad_2d,
                                self.observation_count)


def gradients_struct(planes, sh, viewspace_grad_2d=None, observation_count=None):
    """LumenExGradients over (11, N) gradient planes and (N, 48) SH gradients;
    the view-space accumulators may be left out (NULL)."""
    from ._lib import LumenExGradients
    g = LumenExGradients()
    n = planes.shape[1]
    base = planes.data_ptr()
    for k, name in enumerate(PLANE_NAMES):
        setattr(g, name, base + 4 * k * n)
    g.sh = sh.data_ptr()
    g.viewspace_grad_2d = viewspace_grad_2d.data_ptr() if viewspace_grad_2d is not None else None
    g.observation_count = observation_count.data_ptr() if observation_count is not None else None
    return g


def new_gradients(count: int, device, viewspace: Optional["GaussianGradients"] = None):
    """Uninitialised gradient planes for a backward that writes them
    (LUMEN_EX_GRAD_ASSIGN), with the caller's view-space accumulators if any:
    (planes, sh, LumenExGradients)."""
    planes = torch.empty((11, count), dtype=torch.float32, device=device)
    sh = torch.empty((count, 48), dtype=torch.float32, device=device)
    vs = viewspace.viewspace_grad_2d if viewspace is not None else None
    cnt = viewspace.observation_count if viewspace is not None else None
    return planes, sh, gradients_struct(planes, sh, vs, cnt)


class RasterizeSplats(torch.autograd.Function):
    """(planes (11, N), sh (N, 48)) -> (rgb (H, W, 3), depth (H, W))."""

    @staticmethod
    def forward(ctx, planes, sh, camera: PinholeCamera, config: RenderConfig,
                rasterizer: Rasterizer, viewspace: Optional[GaussianGradients]):
        # an unused output's gradient (the depth's, usually) arrives as None,
        # not as a zero-filled image
        ctx.set_materialize_grads(False)
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
        # the backward writes every gradient entry (no zeroed planes needed);
        # densification statistics accumulate across views in the caller's object
        planes, sh, gs = new_gradients(ctx.scene.count, ctx.scene.device, ctx.viewspace)
        g_rgb = grad_rgb.contiguous() if grad_rgb is not None else \
            torch.zeros((ctx.camera.config.height, ctx.camera.config.width, 3),
                        device=ctx.scene.device)
        g_d = grad_depth.contiguous() if grad_depth is not None else None
        R.render_backward(ctx.scene, ctx.camera, ctx.config, g_rgb, g_d, gs, assign=True)
        return planes, sh, None, None, None, None


def rasterize(planes: torch.Tensor, sh: torch.Tensor, camera: PinholeCamera,
              config: Optional[RenderConfig] = None, rasterizer: Optional[Rasterizer] = None,
              viewspace: Optional[GaussianGradients] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """Differentiable rasterize_scene.  If `viewspace` is given, its
    view-space gradient accumulator (and parameter gradients) also receive
    this view's contribution during backward."""
    return RasterizeSplats.apply(planes, sh, camera, config or RenderConfig(),
                                 rasterizer or default_rasterizer(), viewspace)


# ---- losses: src/optim/loss.mojo ------------------------------------------

class PhotometricLoss(torch.autograd.Function):
    """(pred, target) -> ((1 - lambda) L1 + lambda (1 - SSIM), L1, SSIM) through
    lumen_ex_photometric_loss (one HIP pass computes the values and
    d(loss)/d(pred)); only the first output is differentiable."""

    @staticmethod
    def forward(ctx, pred, target, lambda_dssim: float, rasterizer: Optional[Rasterizer]):
        import ctypes
        from . import _lib
        if not pred.is_cuda or pred.dim() != 3 or pred.shape[2] != 3:
            raise ValueError("photometric loss: pred must be a (H, W, 3) device tensor")
        if target.shape != pred.shape:
            raise ValueError("photometric loss: target shape differs from pred")
        R = rasterizer or default_rasterizer()
        p = pred.detach().float().contiguous()
        t = target.detach().float().contiguous()
        H, W = p.shape[0], p.shape[1]
        out = torch.empty(3, dtype=torch.float32, device=p.device)
        grad = torch.empty_like(p) if ctx.needs_input_grad[0] else None
        st = _lib.lib().lumen_ex_photometric_loss(
            R.handle, ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(t.data_ptr()), W, H,
            float(lambda_dssim), ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(grad.data_ptr()) if grad is not None else None,
            ctypes.c_void_p(torch.cuda.current_stream(p.device).cuda_stream))
        _lib.check("lumen_ex_photometric_loss", st)
        ctx.save_for_backward(grad if grad is not None else out)
        ctx.has_grad = grad is not None
        # the three values are views of `out` (no copies); the gradients of
        # L1 and SSIM are never needed, so they are not materialized either
        ctx.set_materialize_grads(False)
        l1, ssim = out[0], out[1]
        ctx.mark_non_differentiable(l1, ssim)
        return out[2], l1, ssim

    @staticmethod
    def backward(ctx, g_loss, g_l1, g_ssim):
        if not ctx.has_grad or g_loss is None:
            return None, None, None, None
        (grad,) = ctx.saved_tensors
        return grad * g_loss, None, None, None


def photometric_loss(pred: torch.Tensor, target: torch.Tensor, lambda_dssim: float = 0.2,
                     rasterizer: Optional[Rasterizer] = None):
    """((1 - lambda) L1 + lambda (1 - SSIM), L1, SSIM) of (H, W, 3) device images."""
    return PhotometricLoss.apply(pred, target, float(lambda_dssim), rasterizer)


def l1_loss(pred: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Mean |pred - target| whose gradient is +1/N where pred >= target and
    -1/N elsewhere (compute_l1_loss, loss.mojo:11-35)."""
    return photometric_loss(pred, target, 0.0)[0]


def ssim_window(pred: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Mean SSIM over non-overlapping 8x8 windows with the three colour
    channels pooled (compute_ssim_window, loss.mojo:65-133); partial windows
    at the right/bottom edges are not counted."""
    return 1.0 - photometric_loss(pred, target, 1.0)[0]


def combined_loss(pred: torch.Tensor, target: torch.Tensor, lambda_dssim: float = 0.2,
                  rasterizer: Optional[Rasterizer] = None) -> torch.Tensor:
    """(1 - lambda) * L1 + lambda * (1 - SSIM) (compute_combined_loss, loss.mojo:136-144)."""
    return photometric_loss(pred, target, lambda_dssim, rasterizer)[0]


def psnr(pred: torch.Tensor, target: torch.Tensor) -> float:
    mse = float(((pred - target) ** 2).mean())
    if mse < 1e-10:
        return 99.99
    import math
    return -4.3429448 * math.log(mse)
