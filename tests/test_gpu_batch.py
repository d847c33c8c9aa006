"""Multi-view frames (lumen_ex_render_batch), hipGraph replay, batched
backward and the autotuner.

Every view of a batched frame must equal a single-view render of the same
camera bit for bit (images, depth, per-view visible/instance counts), which
the single-view parity tests in turn pin to the oracle.
"""
import math
import sys

import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, look_at, orbit
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import random_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def renderer():
    from lumen_amd.batch import MultiViewRenderer
    return MultiViewRenderer(0)


def mixed_cameras():
    cams = []
    for k, (w, h, f) in enumerate([(320, 240, 300.0), (640, 360, 500.0), (256, 256, 260.0),
                                   (480, 272, 420.0)]):
        ang = 2 * math.pi * k / 4
        cfg = CameraConfig(w, h, f, f, w / 2, h / 2, 0.01, 1000.0)
        cams.append(look_at(cfg, (4.5 * math.cos(ang), 0.8, 4.5 * math.sin(ang))))
    return cams


def test_batch_views_equal_single_view(renderer, rasterizer, oracle):
    import torch
    scene = random_scene(60_000, 1.0, 91)
    dev = scene.to_device()
    cams = mixed_cameras()
    cfg = RenderConfig()
    say("[batch] 4 mixed-resolution views")
    out = renderer.render(dev, cams, cfg, want_stats=True)
    renderer.synchronize()
    for k, cam in enumerate(cams):
        single = rasterizer.render(dev, cam, cfg)
        torch.cuda.synchronize()
        assert out.stats[k].visible_splats == single.visible_splats
        assert out.stats[k].total_splat_instances == single.total_splat_instances
        assert torch.equal(out.views_rgb[k], single.rgb), k
        assert torch.equal(out.views_depth[k], single.depth), k
    ref = oracle.rasterize(scene, cams[1], cfg)
    np.testing.assert_array_equal(out.views_rgb[1].cpu().numpy(), ref["rgb"])


def test_graph_replays_match_eager(renderer):
    """Replays of the captured frame, with eager work and synchronizations in
    between, each equal the eager (plain) frame."""
    import torch
    scene = random_scene(80_000, 1.0, 92)
    dev = scene.to_device()
    cams = orbit(CameraConfig(400, 300, 380.0, 380.0, 200.0, 150.0), 6, 5.0)
    cfg = RenderConfig()
    eager = renderer.render(dev, cams, cfg, graph=False)
    renderer.synchronize()
    ref_rgb, ref_depth = eager.rgb.clone(), eager.depth.clone()
    for rep in range(4):
        say(f"[batch] graph replay {rep}")
        out = renderer.render(dev, cams, cfg, graph=True)
        st = renderer.status(len(cams))
        assert sum(s.visible_splats for s in st) > 0
        junk = torch.randn(1000, 1000, device="cuda") @ torch.randn(1000, 1000, device="cuda")
        torch.cuda.synchronize()
        assert torch.equal(out.rgb, ref_rgb), rep
        assert torch.equal(out.depth, ref_depth), rep
        del junk


def test_single_view_graph_with_sort_pass_in_emit(renderer, oracle):
    """A one-view frame of 3600 tiles captured as a graph: the key emission
    runs the sort's first pass inside the graph (digit table, column scans
    in the planning kernel).  Replays, with eager work and synchronizations
    in between, equal the eager frame and the oracle; a replay whose scene
    outgrew the reserved keys reports the capacity error without faulting."""
    import torch
    from lumen_amd._lib import LumenError
    scene = random_scene(100_000, 1.0, 95)
    dev = scene.to_device()
    cam = look_at(CameraConfig(1280, 720, 1000.0, 1000.0, 640.0, 360.0), (0.4, 0.3, 5.0))
    cfg = RenderConfig()
    eager = renderer.render(dev, [cam], cfg, graph=False)
    renderer.synchronize()
    ref_rgb = eager.rgb.clone()
    ref = oracle.rasterize(scene, cam, cfg)
    np.testing.assert_array_equal(ref_rgb.cpu().numpy().reshape(ref["rgb"].shape), ref["rgb"])
    for rep in range(3):
        say(f"[batch] single-view graph replay {rep}")
        out = renderer.render(dev, [cam], cfg, graph=True)
        renderer.status(1)
        junk = torch.randn(1000, 1000, device="cuda") @ torch.randn(1000, 1000, device="cuda")
        torch.cuda.synchronize()
        assert torch.equal(out.rgb, ref_rgb), rep
        del junk
    dev.planes[3:6] *= 8.0
    out = renderer.render(dev, [cam], cfg, graph=True)
    with pytest.raises(LumenError) as ei:
        renderer.status(1)
    assert ei.value.status == -101
    renderer.synchronize()
    assert float(out.rgb.abs().max()) == 0.0


def test_batch_sorted_stages_without_kept_keys(renderer, rasterizer, oracle):
    """Three 1280x720 views (3600 tiles each) in one frame that does not keep
    its unsorted keys: each view's sorted keys, values and tile ranges equal a
    single-view render that kept them, as do the images; graph replays equal
    the eager frame, and one view equals the oracle."""
    import torch
    scene = random_scene(100_000, 1.0, 96)
    dev = scene.to_device()
    cfg = RenderConfig()
    ccfg = CameraConfig(1280, 720, 1000.0, 1000.0, 640.0, 360.0)
    cams = [look_at(ccfg, eye) for eye in ((0.4, 0.3, 5.0), (4.0, 1.0, 2.5), (-3.0, -0.5, -3.8))]
    say("[batch] 3 views of 1280x720, sorted stages")
    out = renderer.render(dev, cams, cfg, want_stats=True)
    renderer.synchronize()
    st = out.stats
    M = sum(s.total_splat_instances for s in st)
    T = 3 * 80 * 45
    R = renderer.rasterizer
    keys = R.frame_stage("keys", M, np.uint64)
    vals = R.frame_stage("vals", M, np.uint32)
    ranges = R.frame_stage("ranges", 2 * T, np.int32).reshape(T, 2)
    ref_rgb = out.rgb.clone()
    tile_base = 0
    for k, cam in enumerate(cams):
        one = rasterizer.render(dev, cam, cfg, keep_stages=True)
        torch.cuda.synchronize()
        assert one.emit_passes == 0
        kb, n, vb = st[k].key_base, st[k].total_splat_instances, st[k].vis_base
        assert n == one.total_splat_instances and n > 0
        assert torch.equal(out.views_rgb[k], one.rgb), k
        assert torch.equal(out.views_depth[k], one.depth), k
        np.testing.assert_array_equal(keys[kb:kb + n] - (np.uint64(tile_base) << np.uint64(32)),
                                      rasterizer.stage("keys"))
        np.testing.assert_array_equal(vals[kb:kb + n] - vb, rasterizer.stage("vals"))
        rb = ranges[tile_base:tile_base + 3600].copy()
        rb[rb[:, 0] >= 0] -= kb
        np.testing.assert_array_equal(rb, rasterizer.stage("ranges"))
        tile_base += 3600
    ref = oracle.rasterize(scene, cams[1], cfg)
    np.testing.assert_array_equal(out.views_rgb[1].cpu().numpy().reshape(ref["rgb"].shape),
                                  ref["rgb"])
    for rep in range(3):
        say(f"[batch] 3-view graph replay {rep}")
        g = renderer.render(dev, cams, cfg, graph=True)
        renderer.status(3)
        junk = torch.randn(1000, 1000, device="cuda") @ torch.randn(1000, 1000, device="cuda")
        torch.cuda.synchronize()
        assert torch.equal(g.rgb, ref_rgb), rep
        del junk


def test_graph_capacity_overflow_is_reported(renderer):
    """A replay whose scene grew past the reserved key capacity reports
    LUMEN_EX_ERROR_CAPACITY and writes background, without faulting."""
    import torch
    from lumen_amd._lib import LumenError
    scene = random_scene(20_000, 1.0, 93)
    dev = scene.to_device()
    cams = orbit(CameraConfig(320, 240, 300.0, 300.0, 160.0, 120.0), 2, 5.0)
    cfg = RenderConfig()
    renderer.render(dev, cams, cfg, graph=True)      # plans and captures
    renderer.render(dev, cams, cfg, graph=True)      # replay
    renderer.status(2)
    dev.planes[3:6] *= 8.0                           # far larger splats, same pointers
    out = renderer.render(dev, cams, cfg, graph=True)
    with pytest.raises(LumenError) as ei:
        renderer.status(2)
    assert ei.value.status == -101
    renderer.synchronize()
    assert float(out.rgb.abs().max()) == 0.0


def test_batch_backward_is_sum_of_views(renderer, rasterizer):
    import torch
    from lumen_amd.batch import rasterize_views
    from lumen_amd.training import GaussianGradients, rasterize
    scene = random_scene(30_000, 1.0, 94)
    cams = mixed_cameras()[:3]
    cfg = RenderConfig()
    planes = torch.from_numpy(scene.planes).cuda().requires_grad_(True)
    sh = torch.from_numpy(scene.sh).cuda().requires_grad_(True)
    vs = GaussianGradients(scene.count)
    rgb, depth = rasterize_views(planes, sh, cams, cfg, renderer, viewspace=vs)
    w_rgb = torch.linspace(-1, 1, rgb.numel(), device="cuda")
    w_d = torch.linspace(0.5, -0.5, depth.numel(), device="cuda")
    ((rgb * w_rgb).sum() + (depth * w_d).sum()).backward()
    torch.cuda.synchronize()
    gp_b, gs_b = planes.grad.clone(), sh.grad.clone()
    planes.grad = None
    sh.grad = None
    vs1 = GaussianGradients(scene.count)
    base = 0
    total = 0
    for cam in cams:
        P = cam.config.width * cam.config.height
        r, d = rasterize(planes, sh, cam, cfg, rasterizer=rasterizer, viewspace=vs1)
        total = total + (r.reshape(-1) * w_rgb[3 * base: 3 * (base + P)]).sum() + \
            (d.reshape(-1) * w_d[base: base + P]).sum()
        base += P
    total.backward()
    torch.cuda.synchronize()
    for a, b in ((gp_b, planes.grad), (gs_b, sh.grad)):
        err = float((a - b).norm() / b.norm())
        say(f"[batch] backward rel err {err:.2e}")
        assert err < 1e-5
    assert torch.equal(vs.observation_count, vs1.observation_count)
    assert torch.allclose(vs.viewspace_grad_2d, vs1.viewspace_grad_2d, rtol=1e-5, atol=1e-6)


def test_autotune_configs_agree(renderer):
    import torch
    from lumen_amd.batch import autotune
    scene = random_scene(50_000, 1.0, 95)
    dev = scene.to_device()
    cams = orbit(CameraConfig(320, 192, 300.0, 300.0, 160.0, 96.0), 3, 5.0)
    res = autotune(renderer, dev, cams, RenderConfig(), graph=True, reps=2)
    say(f"[batch] autotune best {res.config.tile_width}x{res.config.tile_height}/"
        f"{res.config.batch_depth} table {res.table}")
    assert len(res.table) == 9
    imgs = []
    for cfg in (RenderConfig(), RenderConfig(tile_width=8, tile_height=8, batch_depth=64),
                RenderConfig(tile_width=32, tile_height=8, batch_depth=128)):
        out = renderer.render(dev, cams, cfg)
        renderer.synchronize()
        imgs.append(out.rgb.clone())
    assert torch.equal(imgs[0], imgs[1]) and torch.equal(imgs[0], imgs[2])


def test_batch_autograd_releases_memory(renderer):
    """Repeated forward+backward through rasterize_views keeps the live device
    memory flat (no reference cycle through the autograd context)."""
    import gc
    import torch
    from lumen_amd.batch import rasterize_views
    scene = random_scene(20_000, 1.0, 96)
    cams = mixed_cameras()[:2]
    planes0 = torch.from_numpy(scene.planes).cuda()
    sh0 = torch.from_numpy(scene.sh).cuda()

    def step(graph):
        planes = planes0.clone().requires_grad_(True)
        sh = sh0.clone().requires_grad_(True)
        rgb, depth = rasterize_views(planes, sh, cams, RenderConfig(), renderer, graph=graph)
        (rgb.sum() + depth.sum()).backward()

    for graph in (False, True):
        step(graph)
        torch.cuda.synchronize()
        gc.collect()
        base = torch.cuda.memory_allocated()
        for _ in range(6):
            step(graph)
        torch.cuda.synchronize()
        gc.collect()
        grown = torch.cuda.memory_allocated() - base
        say(f"[batch] autograd memory growth over 6 steps: {grown} B")
        assert grown < 1 << 20


def test_batch_keys_and_ranges_per_view(renderer, rasterizer):
    """Each view of a batched frame is sorted as its own segment with
    view-local tiles; decoded back, its keys (before and after the sort),
    values and tile ranges equal a single-view render's with the view's key,
    visible and tile offsets removed.  Mixed resolutions, and a view that sees
    nothing (an empty segment) in the middle."""
    from lumen_amd.raster import RenderConfig
    scene = random_scene(60_000, 1.0, 99)
    dev = scene.to_device()
    cams = mixed_cameras()
    away = look_at(CameraConfig(200, 150, 200.0, 200.0, 100.0, 75.0), (0.0, 0.0, 40.0),
                   (0.0, 0.0, 80.0))  # looks away from the cloud
    cams = cams[:2] + [away] + cams[2:]
    cfg = RenderConfig()
    out = renderer.render(dev, cams, cfg, want_stats=True, keep_stages=True)
    renderer.synchronize()
    st = out.stats
    assert st[2].total_splat_instances == 0 and st[2].visible_splats == 0
    R = renderer.rasterizer
    M = sum(s.total_splat_instances for s in st)
    T = sum(cfg.tiles(c.config.width, c.config.height)[0] * cfg.tiles(c.config.width, c.config.height)[1]
            for c in cams)
    keys_u = R.frame_stage("keys_unsorted", M, np.uint64)
    vals_u = R.frame_stage("vals_unsorted", M, np.uint32)
    keys = R.frame_stage("keys", M, np.uint64)
    vals = R.frame_stage("vals", M, np.uint32)
    ranges = R.frame_stage("ranges", 2 * T, np.int32).reshape(T, 2)
    tile_base = 0
    for k, cam in enumerate(cams):
        tx, ty = cfg.tiles(cam.config.width, cam.config.height)
        one = rasterizer.render(dev, cam, cfg, keep_stages=True)
        kb, n, vb = st[k].key_base, st[k].total_splat_instances, st[k].vis_base
        assert n == one.total_splat_instances
        shift = np.uint64(tile_base) << np.uint64(32)
        if n:
            np.testing.assert_array_equal(keys_u[kb:kb + n] - shift, rasterizer.stage("keys_unsorted"))
            np.testing.assert_array_equal(keys[kb:kb + n] - shift, rasterizer.stage("keys"))
            np.testing.assert_array_equal(vals_u[kb:kb + n] - vb, rasterizer.stage("vals_unsorted"))
            np.testing.assert_array_equal(vals[kb:kb + n] - vb, rasterizer.stage("vals"))
        r1 = rasterizer.stage("ranges")
        rb = ranges[tile_base:tile_base + tx * ty].copy()
        rb[rb[:, 0] >= 0] -= kb
        np.testing.assert_array_equal(rb, r1)
        tile_base += tx * ty
    say(f"[batch] per-view keys, values and ranges of {len(cams)} views (one empty) match")


def test_batch_wide_keys_and_many_views(renderer, rasterizer):
    """Two regimes of the per-view sort segments: a frame whose keys need 64
    bits (far depth range) and a frame of 40 small views; every view's image
    and per-view keys equal its single-view render."""
    import torch
    from lumen_amd.raster import RenderConfig
    cfg = RenderConfig()
    wide = random_scene(30_000, 1000.0, 81)
    wide.planes[3:6] *= 1000.0
    small = random_scene(20_000, 1.0, 82)
    cases = [
        (wide, [look_at(CameraConfig(640, 480, 600.0, 600.0, 320.0, 240.0, 0.01, 1.0e6), eye)
                for eye in ((0.0, 0.0, 5000.0), (3000.0, 500.0, 4000.0))], 8),
        (small, orbit(CameraConfig(64, 48, 60.0, 60.0, 32.0, 24.0), 40, 4.0), None),
    ]
    for scene, cams, key_bytes in cases:
        dev = scene.to_device()
        out = renderer.render(dev, cams, cfg, want_stats=True, keep_stages=True)
        renderer.synchronize()
        M = sum(s.total_splat_instances for s in out.stats)
        keys = renderer.rasterizer.frame_stage("keys", M, np.uint64)
        tile_base = 0
        for k, cam in enumerate(cams):
            one = rasterizer.render(dev, cam, cfg, keep_stages=True)
            if key_bytes:
                assert one.key_bytes == key_bytes
            torch.cuda.synchronize()
            assert torch.equal(out.views_rgb[k], one.rgb), k
            kb, n = out.stats[k].key_base, out.stats[k].total_splat_instances
            if n:
                np.testing.assert_array_equal(keys[kb:kb + n] - (np.uint64(tile_base) << np.uint64(32)),
                                              rasterizer.stage("keys"))
            tx, ty = cfg.tiles(cam.config.width, cam.config.height)
            tile_base += tx * ty
        say(f"[batch] {len(cams)} views, {M} keys: per-view images and keys match")


def test_reserved_frames_equal_plain_and_report_overflow(renderer):
    """LUMEN_EX_RESERVED: a frame enqueued with the reserved capacities and no
    host round trip equals the plain frame bit for bit (and its backward
    works); a frame that outgrows the capacity reports LUMEN_EX_ERROR_CAPACITY
    and renders background, and its backward writes zero gradients."""
    import torch
    from lumen_amd._lib import LumenError
    from lumen_amd.batch import MultiViewRenderer, rasterize_views
    scene = random_scene(80_000, 1.0, 88)
    dev = scene.to_device()
    cams = orbit(CameraConfig(400, 300, 380.0, 380.0, 200.0, 150.0), 6, 5.0)
    cfg = RenderConfig()
    plain = renderer.render(dev, cams, cfg, want_stats=True)
    renderer.synchronize()
    M = sum(s.total_splat_instances for s in plain.stats)
    R = MultiViewRenderer(0)
    R.rasterizer.reserve(scene.count, cams, cfg, max_instances=M + M // 4)
    held = R.rasterizer.workspace_bytes()
    for rep in range(3):
        out = R.render(dev, cams, cfg, reserved=True)
        st = R.status(len(cams))
        assert [s.total_splat_instances for s in st] == [s.total_splat_instances for s in plain.stats]
        assert torch.equal(out.rgb, plain.rgb) and torch.equal(out.depth, plain.depth), rep
    assert R.rasterizer.workspace_bytes() == held  # nothing allocated
    # backward after a reserved frame equals the plain one's
    grads = []
    for reserved in (False, True):
        planes = dev.planes.detach().clone().requires_grad_(True)
        sh = dev.sh.detach().clone().requires_grad_(True)
        rgb, _ = rasterize_views(planes, sh, cams, cfg, R, reserved=reserved)
        rgb.sum().backward()
        torch.cuda.synchronize()
        grads.append((planes.grad, sh.grad))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])
    # a context sized by a small frame cannot hold the big one
    small = random_scene(2_000, 1.0, 89).to_device()
    R2 = MultiViewRenderer(0)
    R2.render(small, cams[:1], cfg)
    R2.synchronize()
    big = R2.render(dev, cams, cfg, reserved=True)
    with pytest.raises(LumenError) as ei:
        R2.status(len(cams))
    assert ei.value.status == -101
    assert float(big.rgb.abs().max()) == 0.0
    planes = dev.planes.detach().clone().requires_grad_(True)
    sh = dev.sh.detach().clone().requires_grad_(True)
    rgb, _ = rasterize_views(planes, sh, cams, cfg, R2, reserved=True)
    rgb.sum().backward()
    torch.cuda.synchronize()
    assert float(planes.grad.abs().max()) == 0.0 and float(sh.grad.abs().max()) == 0.0
    R.close()
    R2.close()
