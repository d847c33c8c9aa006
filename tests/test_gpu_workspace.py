"""A reserved workspace: lumen_ex_context_reserve allocates exactly what
lumen_ex_workspace_size reports, and frames within its bounds (forward,
backward, photometric loss, single view and a graph-captured batch) add no
device allocation to the context."""
import sys

import pytest

from lumen_amd.camera import CameraConfig, look_at, orbit
from lumen_amd.raster import Rasterizer, RenderConfig, workspace_size
from lumen_amd.scene import perturbed, random_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_reserved_single_view_frames_do_not_allocate():
    import torch
    from lumen_amd.training import combined_loss, rasterize
    scene = random_scene(150_000, 1.0, 97)
    dev = scene.to_device()
    cam = look_at(CameraConfig(1280, 720, 1000.0, 1000.0, 640.0, 360.0), (0.0, 0.0, 5.0))
    cam2 = look_at(CameraConfig(1280, 720, 1000.0, 1000.0, 640.0, 360.0), (0.6, 0.3, 5.2))
    cfg = RenderConfig()
    probe = Rasterizer(0)
    m = max(probe.render(dev, c, cfg).total_splat_instances for c in (cam, cam2))
    probe.close()
    want = workspace_size(scene.count, [cam], cfg, max_instances=m, backward=True)
    R = Rasterizer(0)
    assert R.workspace_bytes() == 0
    R.reserve(scene.count, cam, cfg, max_instances=m, backward=True)
    held = R.workspace_bytes()
    say(f"[workspace] M={m}: reserved {held} B, reported {want} B")
    assert held == want
    target = R.render(perturbed(scene, 0.01, 7).to_device(), cam, cfg).rgb.clone()
    planes = dev.planes.clone().requires_grad_(True)
    sh = dev.sh.clone().requires_grad_(True)
    for c in (cam, cam2, cam):
        img, _ = rasterize(planes, sh, c, cfg, rasterizer=R)
        combined_loss(img, target, rasterizer=R).backward()
        torch.cuda.synchronize()
        assert R.workspace_bytes() == held
    R.close()


def test_reserved_graph_batch_does_not_allocate():
    import torch
    from lumen_amd.batch import MultiViewRenderer
    scene = random_scene(100_000, 1.0, 98)
    dev = scene.to_device()
    cams = orbit(CameraConfig(640, 360, 500.0, 500.0, 320.0, 180.0), 4, 5.0)
    cfg = RenderConfig()
    probe = MultiViewRenderer(0)
    probe.render(dev, cams, cfg, want_stats=True)
    m = sum(s.total_splat_instances for s in probe.status(len(cams)))
    probe.close()
    R = Rasterizer(0)
    R.reserve(scene.count, cams, cfg, max_instances=m, save_pixels=True, graph=True)
    held = R.workspace_bytes()
    assert held == workspace_size(scene.count, cams, cfg, max_instances=m, save_pixels=True,
                                  graph=True)
    mv = MultiViewRenderer(0, rasterizer=R)
    outs = []
    for rep in range(3):
        out = mv.render(dev, cams, cfg, graph=True)
        mv.synchronize()
        outs.append(out.rgb.clone())
        assert R.workspace_bytes() == held, rep
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])
    say(f"[workspace] batch of {len(cams)} views, M={m}: {held} B held throughout")
    mv.close()
