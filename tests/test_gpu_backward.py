"""GPU backward (lumen_ex_render_backward via torch.autograd) vs the oracle.

The oracle's FP64 analytic backward is pinned to finite differences in
tests/test_backward_oracle.py; here the FP32 GPU gradients are compared with
it (relative L2 error per parameter group) and, on a 2k-splat scene, directly
with FP64 central differences of the oracle's forward.
"""
import math
import sys

import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, look_at
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import GaussianScene, random_scene

pytestmark = pytest.mark.gpu

GROUPS = {"pos": slice(0, 3), "scale": slice(3, 6), "rot": slice(6, 10), "opacity": slice(10, 11)}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def subset_scene(n=2000, seed=31):
    rng = np.random.default_rng(seed)
    pos = rng.normal(0, 0.6, (n, 3))
    scale = np.exp(rng.normal(math.log(0.03), 0.4, (n, 3)))
    rot = rng.standard_normal((n, 4))
    op = 1 / (1 + np.exp(-rng.normal(0, 1.2, n)))
    sh = rng.normal(0, 0.08, (n, 48))
    sh[:, :3] = rng.normal(0.2, 0.4, (n, 3))
    return GaussianScene.from_arrays(pos, scale, rot, op, sh)


def gpu_grads(scene, cam, cfg, g_rgb, g_d):
    import torch
    from lumen_amd.training import GaussianGradients, rasterize
    planes = torch.from_numpy(scene.planes).cuda().requires_grad_(True)
    sh = torch.from_numpy(scene.sh).cuda().requires_grad_(True)
    vs = GaussianGradients(scene.count)
    rgb, depth = rasterize(planes, sh, cam, cfg, viewspace=vs)
    loss = (rgb * torch.from_numpy(g_rgb).cuda().float()).sum() + \
           (depth * torch.from_numpy(g_d).cuda().float()).sum()
    loss.backward()
    torch.cuda.synchronize()
    return (planes.grad.cpu().numpy().astype(np.float64), sh.grad.cpu().numpy().astype(np.float64),
            vs.viewspace_grad_2d.cpu().numpy(), vs.observation_count.cpu().numpy())


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("white,depth_on", [(False, True), (True, False)])
def test_backward_matches_fp64_oracle(oracle, white, depth_on):
    scene = subset_scene()
    cam = look_at(CameraConfig(160, 120, 150.0, 150.0, 80.0, 60.0, 0.01, 1000.0), (0.3, -0.2, 3.0))
    cfg = RenderConfig(white_background=white, compute_depth=depth_on)
    rng = np.random.default_rng(1)
    g_rgb = rng.normal(0, 1, (120, 160, 3))
    g_d = rng.normal(0, 0.2, (120, 160))
    say(f"[bwd] 2k subset white={white} depth={depth_on}")
    gp, gs, vs, cnt = gpu_grads(scene, cam, cfg, g_rgb, g_d)
    ref = oracle.backward64(scene.planes, scene.sh, cam, cfg, g_rgb, g_d if depth_on else 0 * g_d)
    for name, sl in GROUPS.items():
        e = rel_err(gp[sl], ref["planes"][sl])
        say(f"[bwd]   {name}: rel L2 err {e:.2e}")
        assert e < 2e-3, name
    e = rel_err(gs, ref["sh"])
    assert e < 2e-3
    np.testing.assert_array_equal(cnt, ref["count"])
    assert rel_err(vs, ref["viewspace"]) < 2e-3


def test_backward_thin_rotated_splats(oracle):
    """Needle-like splats at every angle (axis ratios 20:1 to 1000:1, the
    2D conics' conditioning around the bound up to which the compositors test
    each wave's strip against the ellipse exactly): the backward's 4-row
    strips drop no splat that reaches a pixel, so the gradients stay the
    FP64 oracle's."""
    rng = np.random.default_rng(41)
    n = 1500
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    pos[:, 2] *= 0.3
    long_ = np.exp(rng.uniform(math.log(0.05), math.log(1.2), n))
    ratio = np.exp(rng.uniform(math.log(20.0), math.log(1000.0), n))
    scale = np.stack([long_, long_ / ratio, long_ / ratio], axis=1)
    ang = rng.uniform(0, math.pi, n)
    rot = np.stack([np.cos(ang / 2), np.zeros(n), np.zeros(n), np.sin(ang / 2)], axis=1)
    rot[n // 2:] = rng.standard_normal((n - n // 2, 4))
    op = np.exp(rng.uniform(math.log(1.0 / 255.0), 0.0, n)).clip(0.0, 0.99)
    sh = rng.normal(0, 0.08, (n, 48))
    sh[:, :3] = rng.normal(0.2, 0.4, (n, 3))
    scene = GaussianScene.from_arrays(pos, scale, rot, op, sh)
    cam = look_at(CameraConfig(160, 120, 150.0, 150.0, 80.0, 60.0, 0.01, 1000.0), (0.0, 0.0, 3.0))
    cfg = RenderConfig()
    g = np.random.default_rng(2)
    g_rgb = g.normal(0, 1, (120, 160, 3))
    g_d = g.normal(0, 0.2, (120, 160))
    say("[bwd] thin rotated splats")
    gp, gs, vs, cnt = gpu_grads(scene, cam, cfg, g_rgb, g_d)
    ref = oracle.backward64(scene.planes, scene.sh, cam, cfg, g_rgb, g_d)
    for name, sl in GROUPS.items():
        e = rel_err(gp[sl], ref["planes"][sl])
        say(f"[bwd]   {name}: rel L2 err {e:.2e}")
        assert e < 2e-3, name
    assert rel_err(gs, ref["sh"]) < 2e-3
    np.testing.assert_array_equal(cnt, ref["count"])


def test_backward_against_fp64_finite_differences(oracle):
    """GPU gradient vs central differences of the FP64 forward, on a sample of
    parameters of a 2k-splat scene; samples whose renders change a discrete
    decision are excluded."""
    scene = subset_scene(2000, 44)
    cam = look_at(CameraConfig(128, 96, 120.0, 120.0, 64.0, 48.0, 0.01, 1000.0), (0.2, 0.1, 3.0))
    cfg = RenderConfig()
    rng = np.random.default_rng(2)
    g_rgb = rng.normal(0, 1, (96, 128, 3))
    g_d = rng.normal(0, 0.2, (96, 128))
    gp, gs, _, cnt = gpu_grads(scene, cam, cfg, g_rgb, g_d)
    planes = scene.planes.astype(np.float64)
    sh = scene.sh.astype(np.float64)
    _, _, sig0 = oracle.forward64(planes, sh, cam, cfg)
    visible = np.nonzero(cnt)[0]
    picks = rng.choice(visible, 60, replace=False)
    scale_ref = np.abs(gp).max()
    checked = 0
    for i in picks:
        for k in list(range(11)) + [11 + int(rng.integers(0, 48))]:
            P, S = planes.copy(), sh.copy()
            if k < 11:
                x0 = P[k, i]
            else:
                x0 = S[i, k - 11]
            h = 1e-6 * max(1.0, abs(x0))
            vals = []
            sigs = []
            for sgn in (1, -1):
                P2, S2 = planes.copy(), sh.copy()
                if k < 11:
                    P2[k, i] += sgn * h
                else:
                    S2[i, k - 11] += sgn * h
                rgb, d, sig = oracle.forward64(P2, S2, cam, cfg)
                vals.append(np.sum(rgb * g_rgb) + np.sum(d * g_d))
                sigs.append(sig)
            if sigs[0] != sig0 or sigs[1] != sig0:
                continue
            fd = (vals[0] - vals[1]) / (2 * h)
            an = gp[k, i] if k < 11 else gs[i, k - 11]
            assert abs(fd - an) <= 5e-3 * abs(fd) + 1e-4 * scale_ref, (i, k, fd, an)
            checked += 1
    say(f"[bwd] FD samples checked: {checked}")
    assert checked >= 600


def test_losses_match_reference_semantics(oracle):
    """l1_loss / ssim_window / combined_loss (torch, on device) vs the oracle's
    restatement of src/optim/loss.mojo, and the L1 subgradient convention."""
    import torch
    from lumen_amd.training import combined_loss, l1_loss, ssim_window
    rng = np.random.default_rng(3)
    a = rng.uniform(0, 1, (40, 56, 3)).astype(np.float32)
    b = np.clip(a + rng.normal(0, 0.1, a.shape), 0, 1).astype(np.float32)
    b[0, 0, 0] = a[0, 0, 0]  # diff == 0 -> +1/N
    ta = torch.from_numpy(a).cuda().requires_grad_(True)
    tb = torch.from_numpy(b).cuda()
    l1 = l1_loss(ta, tb)
    l1.backward()
    g = np.zeros(a.size, np.float32)
    ref_l1 = oracle.lib().orc_l1_loss(oracle.fptr(a.ravel().copy()), oracle.fptr(b.ravel().copy()),
                                      a.size, oracle.fptr(g))
    assert float(l1.detach()) == pytest.approx(ref_l1, rel=1e-5)
    np.testing.assert_allclose(ta.grad.cpu().numpy().ravel(), g, rtol=1e-6)
    s = float(ssim_window(ta.detach(), tb))
    ref_s = oracle.lib().orc_ssim(oracle.fptr(a.ravel().copy()), oracle.fptr(b.ravel().copy()), 56, 40, 8)
    assert s == pytest.approx(ref_s, rel=1e-5)
    c = float(combined_loss(ta.detach(), tb))
    assert c == pytest.approx(oracle.lib().orc_combined_loss(ref_l1, ref_s, 0.2), rel=1e-5)


@pytest.mark.parametrize("H,W", [(40, 56), (37, 61), (1080, 1920), (5, 7)])
def test_photometric_loss_vs_float64_reference(H, W):
    """lumen_ex_photometric_loss (values and gradient, one HIP pass) vs the
    float64 PyTorch formulation of loss.mojo and its autograd gradient; sizes
    with partial windows and with no full window."""
    import torch
    from loss_reference import combined_ref, l1_ref, ssim_ref
    from lumen_amd.training import photometric_loss
    g = torch.Generator().manual_seed(H * 1000 + W)
    a = torch.rand((H, W, 3), generator=g)
    b = (a + 0.1 * torch.randn((H, W, 3), generator=g)).clamp(0, 1)
    b[0, 0, 0] = a[0, 0, 0]  # diff == 0 takes +1/N
    pa = a.cuda().requires_grad_(True)
    loss, l1, ssim = photometric_loss(pa, b.cuda(), 0.2)
    loss.backward()
    ra = a.double().requires_grad_(True)
    ref = combined_ref(ra, b.double())
    ref.backward()
    assert float(l1) == pytest.approx(float(l1_ref(a.double(), b.double())), rel=2e-6)
    assert float(ssim) == pytest.approx(float(ssim_ref(a.double(), b.double())), rel=2e-5, abs=1e-6)
    assert float(loss.detach()) == pytest.approx(float(ref.detach()), rel=2e-5)
    got = pa.grad.cpu().double()
    err = float((got - ra.grad).norm() / ra.grad.norm())
    say(f"[loss] {H}x{W} rel grad err {err:.2e}")
    assert err < 1e-4


def test_fwd_bwd_1m_1080p_descends(rasterizer):
    """The 1M-splat 1080p training step: L1 + D-SSIM against a render of a
    perturbed copy; gradients finite, and a small step on the means lowers the
    loss."""
    import torch
    from lumen_amd.scene import perturbed, workload
    from lumen_amd.training import combined_loss, rasterize
    scene, cam = workload("random1m_1920x1080")
    target_scene = perturbed(scene, 0.01, 7)
    cfg = RenderConfig()
    with torch.no_grad():
        target = rasterizer.render(target_scene.to_device(), cam, cfg).rgb.clone()
    planes = torch.from_numpy(scene.planes).cuda().requires_grad_(True)
    sh = torch.from_numpy(scene.sh).cuda().requires_grad_(True)
    rgb, _ = rasterize(planes, sh, cam, cfg, rasterizer=rasterizer)
    loss = combined_loss(rgb, target)
    loss.backward()
    torch.cuda.synchronize()
    g = planes.grad
    assert torch.isfinite(g).all() and torch.isfinite(sh.grad).all()
    assert float(g[0:3].abs().sum()) > 0
    with torch.no_grad():
        step = planes.detach().clone()
        gpos = g[0:3]
        step[0:3] -= 1e-4 * gpos / gpos.abs().max()
        rgb2, _ = rasterize(step, sh.detach(), cam, cfg, rasterizer=rasterizer)
        loss2 = combined_loss(rgb2, target)
    say(f"[bwd] 1M loss {float(loss):.6f} -> {float(loss2):.6f}")
    assert float(loss2) < float(loss)


def test_backward_is_deterministic(rasterizer):
    """One view: every partial gradient has its own slot and is summed in a
    fixed order, so repeated backward passes agree bit for bit."""
    import torch
    from lumen_amd.training import GaussianGradients, rasterize
    scene = random_scene(100_000, 1.0, 41)
    cfg_cam = CameraConfig(640, 480, 600.0, 600.0, 320.0, 240.0)
    cam = look_at(cfg_cam, (0.0, 0.5, 4.5))
    w = None
    results = []
    for rep in range(3):
        planes = torch.from_numpy(scene.planes).cuda().requires_grad_(True)
        sh = torch.from_numpy(scene.sh).cuda().requires_grad_(True)
        vs = GaussianGradients(scene.count)
        rgb, depth = rasterize(planes, sh, cam, RenderConfig(), rasterizer=rasterizer, viewspace=vs)
        if w is None:
            g = torch.Generator(device="cuda").manual_seed(3)
            w = torch.randn(rgb.shape, generator=g, device="cuda")
        (rgb * w).sum().backward()
        torch.cuda.synchronize()
        results.append((planes.grad.clone(), sh.grad.clone(), vs.viewspace_grad_2d.clone()))
        say(f"[bwd] determinism rep {rep}")
    for r in results[1:]:
        for a, b in zip(results[0], r):
            assert torch.equal(a, b)


@pytest.mark.parametrize("views", [1, 3])
def test_assign_equals_accumulate_into_zeros(rasterizer, views):
    """LUMEN_EX_GRAD_ASSIGN writes exactly what accumulating into zeroed planes
    gives, over planes full of garbage, including the zeros of Gaussians no
    view sees (one view, and a batch of three); the view-space statistics
    still accumulate."""
    import torch
    from lumen_amd.batch import MultiViewRenderer
    from lumen_amd.camera import orbit
    from lumen_amd.training import GaussianGradients, gradients_struct
    scene = random_scene(40_000, 1.0, 97)
    dev = scene.to_device()
    cfg = RenderConfig()
    rng = np.random.default_rng(8)
    if views == 1:
        cams = [look_at(CameraConfig(320, 240, 300.0, 300.0, 160.0, 120.0), (0.0, 0.5, 4.0))]
        rasterizer.render(dev, cams[0], cfg, save_pixels=True)
        run = lambda g, assign: rasterizer.render_backward(dev, cams[0], cfg, g_rgb, None, g,  # noqa
                                                          assign=assign)
    else:
        cams = orbit(CameraConfig(320, 240, 300.0, 300.0, 160.0, 120.0), views, 4.5)
        mv = MultiViewRenderer(0, rasterizer=rasterizer)
        mv.render(dev, cams, cfg, save_pixels=True)
        mv.synchronize()
        run = lambda g, assign: (mv.backward(dev, cams, cfg, g_rgb, None, g, assign=assign),  # noqa
                                 mv.synchronize())
    P = sum(c.config.width * c.config.height for c in cams)
    g_rgb = torch.from_numpy(rng.normal(0, 1e-3, 3 * P).astype(np.float32)).cuda()
    if views == 1:
        g_rgb = g_rgb.view(cams[0].config.height, cams[0].config.width, 3)
    ref = GaussianGradients(scene.count)
    run(ref, False)
    planes = torch.full((11, scene.count), float("nan"), device="cuda")
    sh = torch.full((scene.count, 48), 1234.5, device="cuda")
    vs = GaussianGradients(scene.count)
    vs.viewspace_grad_2d.fill_(1.0)
    run(gradients_struct(planes, sh, vs.viewspace_grad_2d, vs.observation_count), True)
    torch.cuda.synchronize()
    assert torch.equal(planes, ref.planes) and torch.equal(sh, ref.sh)
    assert torch.equal(vs.viewspace_grad_2d, ref.viewspace_grad_2d + 1.0)
    assert torch.equal(vs.observation_count, ref.observation_count)
    assert int((ref.observation_count == 0).sum()) > 0   # some Gaussians unseen
    say(f"[bwd] assign == accumulate into zeros ({views} view(s))")


@pytest.mark.parametrize("views", [1, 2])
def test_assign_writes_zeros_when_nothing_is_visible(rasterizer, views):
    """A frame in which no splat reaches a tile (the camera looks away from
    the cloud) has no gradient.  With LUMEN_EX_GRAD_ASSIGN the planes come
    from torch.empty, so the backward must still write every entry: planes
    and SH full of NaN / garbage come back all zero."""
    import torch
    from lumen_amd.batch import MultiViewRenderer
    from lumen_amd.training import GaussianGradients, gradients_struct
    scene = random_scene(5_000, 1.0, 12)
    dev = scene.to_device()
    cfg = RenderConfig()
    # behind the camera: every splat fails the depth cull
    cams = [look_at(CameraConfig(160, 120, 150.0, 150.0, 80.0, 60.0), (0.0, 0.0, 6.0 + k),
                    target=(0.0, 0.0, 20.0)) for k in range(views)]
    P = sum(c.config.width * c.config.height for c in cams)
    g_rgb = torch.ones(3 * P, device="cuda")
    planes = torch.full((11, scene.count), float("nan"), device="cuda")
    sh = torch.full((scene.count, 48), float("nan"), device="cuda")
    vs = GaussianGradients(scene.count)
    g = gradients_struct(planes, sh, vs.viewspace_grad_2d, vs.observation_count)
    if views == 1:
        out = rasterizer.render(dev, cams[0], cfg, save_pixels=True)
        assert out.visible_splats == 0 and out.total_splat_instances == 0
        rasterizer.render_backward(dev, cams[0], cfg, g_rgb.view(120, 160, 3), None, g, assign=True)
    else:
        mv = MultiViewRenderer(0, rasterizer=rasterizer)
        mv.render(dev, cams, cfg, save_pixels=True)
        mv.synchronize()
        mv.backward(dev, cams, cfg, g_rgb, None, g, assign=True)
        mv.synchronize()
    torch.cuda.synchronize()
    assert bool((planes == 0).all()) and bool((sh == 0).all())
    assert int(vs.observation_count.sum()) == 0
    say(f"[bwd] nothing visible ({views} view(s)): assign wrote zeros")
