"""Plain PyTorch float64 formulation of src/optim/loss.mojo (compute_l1_loss
:11-35, compute_ssim_window :65-133, compute_combined_loss :136-144), used by
the GPU tests as the reference for lumen_ex_photometric_loss's values and,
through autograd, its gradient."""
import torch

SSIM_C1 = 0.0001  # loss.mojo:8
SSIM_C2 = 0.0009  # loss.mojo:9


def l1_ref(pred, target):
    d = pred - target
    return torch.where(d >= 0, d, -d).mean()


def ssim_ref(pred, target, win=8):
    H, W = pred.shape[0], pred.shape[1]
    hy, wx = H // win, W // win
    if hy == 0 or wx == 0:
        return torch.ones((), dtype=pred.dtype, device=pred.device)

    def blocks(x):
        x = x[: hy * win, : wx * win]
        return x.reshape(hy, win, wx, win, 3).permute(0, 2, 1, 3, 4).reshape(hy, wx, win * win * 3)

    x, y = blocks(pred), blocks(target)
    mx, my = x.mean(-1, keepdim=True), y.mean(-1, keepdim=True)
    dx, dy = x - mx, y - my
    vx, vy, cxy = (dx * dx).mean(-1), (dy * dy).mean(-1), (dx * dy).mean(-1)
    mx, my = mx[..., 0], my[..., 0]
    num = (2.0 * mx * my + SSIM_C1) * (2.0 * cxy + SSIM_C2)
    den = (mx * mx + my * my + SSIM_C1) * (vx + vy + SSIM_C2)
    return (num / den).mean()


def combined_ref(pred, target, lam=0.2):
    return (1.0 - lam) * l1_ref(pred, target) + lam * (1.0 - ssim_ref(pred, target))
