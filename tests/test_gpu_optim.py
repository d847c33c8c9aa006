"""lumen_ex_adam_step, lumen_ex_densify_and_prune and
lumen_ex_reset_opacities on the GPU against the oracle, bit for bit, then a
few steps of the training loop (TrainingCoordinator) end to end."""
import ctypes
import sys

import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, look_at
from lumen_amd.scene import perturbed, random_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grads_like(n, rng, scale=1e-3):
    import torch
    from lumen_amd.training import GaussianGradients
    g = GaussianGradients(n)
    g.planes.copy_(torch.from_numpy(rng.normal(0, scale, (11, n)).astype(np.float32)))
    g.sh.copy_(torch.from_numpy(rng.normal(0, scale, (n, 48)).astype(np.float32)))
    return g


@pytest.mark.parametrize("all_params", [False, True])
def test_adam_step_matches_oracle(oracle, all_params):
    import torch
    from lumen_amd.optim import AdamOptimizer, TrainingConfig
    rng = np.random.default_rng(21)
    n = 100_003
    sc = random_scene(n, 1.0, 83)
    dev = sc.to_device()
    cfg = TrainingConfig(iterations=40)
    opt = AdamOptimizer(n, all_params=all_params)
    planes, sh = sc.planes.copy(), sc.sh.copy()
    m = (np.zeros_like(planes), np.zeros_like(sh))
    v = (np.zeros_like(planes), np.zeros_like(sh))
    ocfg = oracle.train_config(cfg)
    for t in range(1, 4):
        g = _grads_like(n, rng)
        say(f"[optim] adam step {t} all_params={all_params}")
        opt.step(dev, g, cfg)
        oracle.adam_step(planes, sh, g.planes.cpu().numpy(), g.sh.cpu().numpy(), m, v, ocfg, t,
                         all_params)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dev.planes.cpu().numpy(), planes)
        np.testing.assert_array_equal(dev.sh.cpu().numpy(), sh)
        np.testing.assert_array_equal(opt.m_planes.cpu().numpy()[[0, 1, 2, 10]], m[0][[0, 1, 2, 10]])
        np.testing.assert_array_equal(opt.v_planes.cpu().numpy()[[0, 1, 2, 10]], v[0][[0, 1, 2, 10]])
        if all_params:
            np.testing.assert_array_equal(opt.m_sh.cpu().numpy(), m[1])
            np.testing.assert_array_equal(opt.v_planes.cpu().numpy(), v[0])


def _mixed_scene(n, seed):
    rng = np.random.default_rng(seed)
    sc = random_scene(n, 1.0, seed)
    sc.planes[3:6] *= rng.choice([1.0, 30.0, 400.0], size=n)
    sc.planes[6:10] *= 3.0
    vs = rng.exponential(0.0003, n).astype(np.float32)
    cnt = rng.integers(0, 4, n).astype(np.int32)
    return sc, vs, cnt


@pytest.mark.parametrize("n", [1, 255, 257, 70_001])
def test_densify_matches_oracle(oracle, n):
    import torch
    from lumen_amd.optim import TrainingConfig, densify_and_prune
    from lumen_amd.training import GaussianGradients
    sc, vs, cnt = _mixed_scene(n, 84 + n)
    dev = sc.to_device()
    g = GaussianGradients(n)
    g.viewspace_grad_2d.copy_(torch.from_numpy(vs))
    g.observation_count.copy_(torch.from_numpy(cnt))
    cfg = TrainingConfig()
    say(f"[optim] densify n={n}")
    new, stats = densify_and_prune(dev, g, cfg, 4.0)
    op, osh, origin, (cl, sp, pr) = oracle.densify_and_prune(sc.planes, sc.sh, vs, cnt,
                                                             oracle.train_config(cfg), 4.0)
    assert (stats.cloned, stats.split, stats.pruned) == (cl, sp, pr)
    assert new.count == op.shape[1]
    np.testing.assert_array_equal(new.planes.cpu().numpy(), op)
    np.testing.assert_array_equal(new.sh.cpu().numpy(), osh)


def test_densify_carries_adam_moments(oracle):
    import torch
    from lumen_amd.optim import AdamOptimizer, TrainingConfig, densify_and_prune
    from lumen_amd.training import GaussianGradients
    n = 40_000
    sc, vs, cnt = _mixed_scene(n, 85)
    dev = sc.to_device()
    g = GaussianGradients(n)
    g.viewspace_grad_2d.copy_(torch.from_numpy(vs))
    g.observation_count.copy_(torch.from_numpy(cnt))
    opt = AdamOptimizer(n, all_params=True)
    rng = np.random.default_rng(4)
    for t in (opt.m_planes, opt.m_sh, opt.v_planes, opt.v_sh):
        t.copy_(torch.from_numpy(rng.random(tuple(t.shape)).astype(np.float32)))
    m0p, m0s = opt.m_planes.cpu().numpy(), opt.m_sh.cpu().numpy()
    v0p = opt.v_planes.cpu().numpy()
    new, _ = densify_and_prune(dev, g, TrainingConfig(), 4.0, optimizer=opt)
    _, _, origin, _ = oracle.densify_and_prune(sc.planes, sc.sh, vs, cnt,
                                               oracle.train_config(), 4.0)
    src, fresh = origin // 2, origin % 2 == 1
    assert opt.count == new.count == origin.size
    mp, ms, vp = opt.m_planes.cpu().numpy(), opt.m_sh.cpu().numpy(), opt.v_planes.cpu().numpy()
    np.testing.assert_array_equal(mp[:, ~fresh], m0p[:, src[~fresh]])
    np.testing.assert_array_equal(ms[~fresh], m0s[src[~fresh]])
    np.testing.assert_array_equal(vp[:, ~fresh], v0p[:, src[~fresh]])
    assert fresh.any() and not mp[:, fresh].any() and not ms[fresh].any() and not vp[:, fresh].any()


def test_density_control_kats_on_gpu():
    """tests/test_density_control.mojo:15-80 through the GPU entry points."""
    import torch
    from lumen_amd.optim import TrainingConfig, densify_and_prune, reset_opacities
    from lumen_amd.scene import DeviceScene
    from lumen_amd.training import GaussianGradients

    def scene(pos, s, op):
        n = len(pos)
        planes = np.concatenate([np.array(pos, np.float32).T, np.full((3, n), s, np.float32),
                                 np.array([[1.0] * n, [0.0] * n, [0.0] * n, [0.0] * n], np.float32),
                                 np.array([op], np.float32)])
        return DeviceScene(torch.from_numpy(planes).cuda(), torch.zeros((n, 48), device="cuda"))

    new, st = densify_and_prune(scene([(0, 0, 0), (1, 1, 1)], 0.1, [0.8, 0.001]),
                                GaussianGradients(2), TrainingConfig(prune_opacity_threshold=0.005),
                                10.0)
    assert st.pruned == 1 and new.count == 1
    g = GaussianGradients(1)
    g.accumulate_viewspace_grad(0, 0.01)
    new, st = densify_and_prune(scene([(0, 0, 0)], 0.005, [0.8]), g, TrainingConfig(), 10.0)
    assert st.cloned == 1 and new.count == 2
    new, st = densify_and_prune(scene([(0, 0, 0)], 0.2, [0.8]), g, TrainingConfig(), 10.0)
    assert st.split == 1 and new.count == 2
    assert abs(float(new.planes[3, 0]) - 0.125) < 1e-4
    s = scene([(0, 0, 0)], 0.1, [0.95])
    reset_opacities(s, 0.01)
    assert abs(float(s.planes[10, 0]) - 0.01) < 1e-5


def test_reset_opacities_matches_oracle(oracle):
    from lumen_amd.optim import reset_opacities
    sc = random_scene(50_000, 1.0, 86)
    dev = sc.to_device()
    reset_opacities(dev, 0.3)
    ref = sc.planes[10].copy()
    oracle.reset_opacities(ref, 0.3)
    np.testing.assert_array_equal(dev.planes[10].cpu().numpy(), ref)


def test_training_loop_reduces_loss():
    """A few iterations of TrainingCoordinator on a scene fitting a render of a
    perturbed copy: the loss falls, densification runs on schedule and the
    optimizer state follows the scene's size."""
    import torch
    from lumen_amd.optim import AdamOptimizer, TrainingConfig, TrainingCoordinator
    from lumen_amd.raster import Rasterizer, RenderConfig
    sc = random_scene(30_000, 1.0, 87)
    cam = look_at(CameraConfig(320, 240, 300.0, 300.0, 160.0, 120.0), (0.0, 0.0, 5.0))
    R = Rasterizer(0)
    target = R.render(perturbed(sc, 0.02, 9).to_device(), cam, RenderConfig()).rgb.clone()
    dev = sc.to_device()
    cfg = TrainingConfig(iterations=30, densify_from_iter=10, densification_interval=10,
                         densify_until_iter=10, opacity_reset_interval=1000,
                         densify_grad_threshold=1e-9, opacity_lr=0.005,
                         position_lr_init=5e-4, position_lr_final=5e-5)
    coord = TrainingCoordinator(cfg, RenderConfig(), R)
    opt = AdamOptimizer(dev.count)
    losses, counts = [], []
    for it in range(1, 31):
        dev, met = coord.train_step(dev, cam, target, opt, it, 4.0)
        losses.append(met.loss)
        counts.append(met.splat_count)
        assert opt.count == dev.count == met.splat_count
    torch.cuda.synchronize()
    say(f"[optim] loss {losses[0]:.5f} -> {losses[-1]:.5f}; splats {counts[0]} -> {counts[-1]}")
    assert np.all(np.isfinite(losses))
    assert losses[8] < losses[0]            # before densification
    assert counts[9] > counts[8]            # densified at iteration 10
    assert losses[-1] < losses[10]          # and it keeps fitting afterwards
    R.close()
