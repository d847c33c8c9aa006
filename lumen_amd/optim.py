"""The optimizer side of training on the GPU: Adam, adaptive density control
and the training step that ties them to the differentiable renderer.

Mirrors src/optim/optimizer.mojo, density_control.mojo and trainer.mojo, and
TrainingConfig (src/core/config.mojo:113-176), with the same names and
defaults; the arithmetic runs in liblumen.so (lumen_ex_adam_step,
lumen_ex_densify_and_prune, lumen_ex_reset_opacities).

TrainingCoordinator.train_step renders through RasterizeSplats, takes the
combined photometric loss and its real gradient (the reference's train_step
injects placeholder gradients proportional to the loss, trainer.mojo:59-65),
steps Adam and densifies / resets opacities on the reference's schedule.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Tuple

import torch

from . import _lib
from .camera import PinholeCamera
from .raster import Rasterizer, RenderConfig, default_rasterizer
from .scene import DeviceScene, SH_COEFFS
from .training import GaussianGradients, combined_loss, psnr, rasterize


@dataclass
class TrainingConfig:
    """src/core/config.mojo:113-176."""
    iterations: int = 30000
    position_lr_init: float = 0.00016
    position_lr_final: float = 0.0000016
    feature_lr: float = 0.0025
    opacity_lr: float = 0.05
    scaling_lr: float = 0.005
    rotation_lr: float = 0.001
    lambda_dssim: float = 0.2
    densify_from_iter: int = 500
    densify_until_iter: int = 15000
    densification_interval: int = 100
    opacity_reset_interval: int = 3000
    densify_grad_threshold: float = 0.0002
    prune_opacity_threshold: float = 0.005
    max_screen_size_ratio: float = 20.0

    def should_densify(self, iteration: int) -> bool:
        if self.densify_from_iter <= iteration <= self.densify_until_iter:
            return iteration % self.densification_interval == 0
        return False

    def should_reset_opacity(self, iteration: int) -> bool:
        if 0 < iteration <= self.densify_until_iter:
            return iteration % self.opacity_reset_interval == 0
        return False

    def ex_struct(self, beta1: float = 0.9, beta2: float = 0.999,
                  epsilon: float = 1e-8) -> _lib.LumenExTrainConfig:
        c = _lib.LumenExTrainConfig()
        c.iterations = self.iterations
        for k in ("position_lr_init", "position_lr_final", "feature_lr", "opacity_lr",
                  "scaling_lr", "rotation_lr", "densify_grad_threshold",
                  "prune_opacity_threshold"):
            setattr(c, k, getattr(self, k))
        c.beta1, c.beta2, c.epsilon = beta1, beta2, epsilon
        return c


def planes_struct(planes: torch.Tensor, sh: torch.Tensor) -> _lib.LumenExPlanes:
    """LumenExPlanes over an (11, N) plane tensor and an (N, 48) SH tensor."""
    assert planes.is_contiguous() and sh.is_contiguous()
    assert planes.dtype == torch.float32 and sh.dtype == torch.float32
    n = planes.shape[1]
    p = _lib.LumenExPlanes()
    base = planes.data_ptr()
    for k, name in enumerate(("pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z",
                              "rot_w", "rot_x", "rot_y", "rot_z", "opacity")):
        setattr(p, name, base + 4 * k * n)
    p.sh = sh.data_ptr()
    return p


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class AdamOptimizer:
    """src/optim/optimizer.mojo:12-170: first and second moments per
    parameter (here for all 59 of a Gaussian, in the GaussianSoA layout, so
    that all_params can use them too), beta1 0.9, beta2 0.999, eps 1e-8."""

    def __init__(self, count: int, device="cuda", all_params: bool = False):
        z = lambda *s: torch.zeros(s, dtype=torch.float32, device=device)  # noqa: E731
        self.m_planes, self.m_sh = z(11, count), z(count, SH_COEFFS)
        self.v_planes, self.v_sh = z(11, count), z(count, SH_COEFFS)
        self.beta1, self.beta2, self.epsilon = 0.9, 0.999, 1e-8
        self.step_count = 0
        self.all_params = all_params

    @property
    def count(self) -> int:
        return int(self.m_planes.shape[1])

    def get_position_lr(self, config: TrainingConfig, iteration: int) -> float:
        """position_lr_init * (final / init) ^ min(1, iteration / iterations)
        in float32 (optimizer.mojo:103-110; numpy's float32 power here, the C
        library's powf inside lumen_ex_adam_step)."""
        import numpy as np
        f = np.float32
        progress = min(f(1.0), f(iteration) / f(config.iterations))
        ratio = f(config.position_lr_final) / f(config.position_lr_init)
        return float(f(config.position_lr_init) * np.power(ratio, progress, dtype=f))

    def step(self, scene: DeviceScene, grads: GaussianGradients, config: TrainingConfig) -> None:
        """One Adam step of the scene's planes, in place (optimizer.mojo:112-170)."""
        if scene.count != self.count or grads.count != self.count:
            raise ValueError("AdamOptimizer.step: scene, gradients and state differ in size")
        self.step_count += 1
        st = _lib.lib().lumen_ex_adam_step(
            ctypes.byref(planes_struct(scene.planes, scene.sh)), ctypes.byref(grads.ex_struct()),
            ctypes.byref(planes_struct(self.m_planes, self.m_sh)),
            ctypes.byref(planes_struct(self.v_planes, self.v_sh)), self.count,
            ctypes.byref(config.ex_struct(self.beta1, self.beta2, self.epsilon)),
            self.step_count, _lib.LUMEN_EX_ADAM_ALL_PARAMS if self.all_params else 0,
            _stream(scene.device))
        _lib.check("lumen_ex_adam_step", st)


@dataclass
class DensityControlStats:
    """density_control.mojo:14-22."""
    cloned: int = 0
    split: int = 0
    pruned: int = 0


def densify_and_prune(scene: DeviceScene, grads: GaussianGradients, config: TrainingConfig,
                      scene_extent: float, optimizer: Optional[AdamOptimizer] = None,
                      rasterizer: Optional[Rasterizer] = None
                      ) -> Tuple[DeviceScene, DensityControlStats]:
    """densify_and_prune (density_control.mojo:25-98): the new scene, and the
    stats.  With an optimizer, its moments follow the Gaussians (zero for the
    new ones) and it is resized in place."""
    R = rasterizer or default_rasterizer()
    n = scene.count
    cap = max(2 * n, 1)
    dev = scene.device
    out_planes = torch.empty((11, cap), dtype=torch.float32, device=dev)
    out_sh = torch.empty((cap, SH_COEFFS), dtype=torch.float32, device=dev)
    mom_in = mom_out = None
    new_state = None
    if optimizer is not None:
        if optimizer.count != n:
            raise ValueError("densify_and_prune: optimizer state differs in size from the scene")
        new_state = [torch.empty((11, cap), dtype=torch.float32, device=dev),
                     torch.empty((cap, SH_COEFFS), dtype=torch.float32, device=dev),
                     torch.empty((11, cap), dtype=torch.float32, device=dev),
                     torch.empty((cap, SH_COEFFS), dtype=torch.float32, device=dev)]
        mom_in = (_lib.LumenExPlanes * 2)(planes_struct(optimizer.m_planes, optimizer.m_sh),
                                          planes_struct(optimizer.v_planes, optimizer.v_sh))
        mom_out = (_lib.LumenExPlanes * 2)(planes_struct(new_state[0], new_state[1]),
                                           planes_struct(new_state[2], new_state[3]))
    stats = _lib.LumenExDensifyStats()
    sc = scene.ex_struct()
    st = _lib.lib().lumen_ex_densify_and_prune(
        R.handle, ctypes.byref(sc), grads.viewspace_grad_2d.data_ptr(),
        grads.observation_count.data_ptr(), ctypes.byref(config.ex_struct()), float(scene_extent),
        ctypes.byref(planes_struct(out_planes, out_sh)), cap, mom_in, mom_out,
        ctypes.byref(stats), _stream(dev))
    _lib.check("lumen_ex_densify_and_prune", st)
    m = int(stats.count)
    # the outputs were written with a plane stride of cap: compact to m
    new_scene = DeviceScene(out_planes[:, :m].contiguous(), out_sh[:m].contiguous())
    if optimizer is not None:
        optimizer.m_planes = new_state[0][:, :m].contiguous()
        optimizer.m_sh = new_state[1][:m].contiguous()
        optimizer.v_planes = new_state[2][:, :m].contiguous()
        optimizer.v_sh = new_state[3][:m].contiguous()
    return new_scene, DensityControlStats(int(stats.cloned), int(stats.split), int(stats.pruned))


def reset_opacities(scene: DeviceScene, max_reset_val: float = 0.01) -> None:
    """reset_opacities (density_control.mojo:101-109), in place."""
    op = scene.planes[10]
    st = _lib.lib().lumen_ex_reset_opacities(op.data_ptr(), scene.count, float(max_reset_val),
                                             _stream(scene.device))
    _lib.check("lumen_ex_reset_opacities", st)


@dataclass
class TrainingMetrics:
    """trainer.mojo:16-27."""
    iteration: int
    loss: float
    psnr: float
    splat_count: int


class TrainingCoordinator:
    """trainer.mojo:30-76 with the real backward."""

    def __init__(self, train_cfg: TrainingConfig, render_cfg: Optional[RenderConfig] = None,
                 rasterizer: Optional[Rasterizer] = None):
        self.config = train_cfg
        self.render_config = render_cfg or RenderConfig()
        self.rasterizer = rasterizer or default_rasterizer()
        # the view-space statistics, accumulated over the steps until the
        # next densification
        self.stats: Optional[GaussianGradients] = None

    def train_step(self, scene: DeviceScene, camera: PinholeCamera, ground_truth: torch.Tensor,
                   optimizer: AdamOptimizer, iteration: int, scene_radius: float
                   ) -> Tuple[DeviceScene, TrainingMetrics]:
        """Forward, loss, backward, Adam, then density control and opacity
        reset when the schedule says so.  Returns the (possibly new) scene."""
        planes = scene.planes.detach().requires_grad_(True)
        sh = scene.sh.detach().requires_grad_(True)
        if self.stats is None or self.stats.count != scene.count:
            z = lambda dt: torch.zeros(scene.count, dtype=dt, device=scene.device)  # noqa: E731
            self.stats = GaussianGradients.wrap(None, None, z(torch.float32), z(torch.int32))
        img, _ = rasterize(planes, sh, camera, self.render_config, rasterizer=self.rasterizer,
                           viewspace=self.stats)
        loss = combined_loss(img, ground_truth, self.config.lambda_dssim, self.rasterizer)
        loss.backward()
        # the backward's gradient planes as they are, with the statistics
        grads = GaussianGradients.wrap(planes.grad, sh.grad, self.stats.viewspace_grad_2d,
                                       self.stats.observation_count)
        loss_v = float(loss.detach())
        psnr_v = psnr(img.detach(), ground_truth)
        optimizer.step(scene, grads, self.config)
        if self.config.should_densify(iteration):
            scene, _ = densify_and_prune(scene, grads, self.config, scene_radius, optimizer,
                                         self.rasterizer)
            self.stats = None
        if self.config.should_reset_opacity(iteration):
            reset_opacities(scene, 0.01)
        return scene, TrainingMetrics(iteration, loss_v, psnr_v, scene.count)
