"""The batched and served workloads at their full sizes.

- Multi-view: 3M Gaussians, 32 cameras on a closed Catmull-Rom orbit of
  radius 6 at 1920x1080, as one multi-view frame (eager, and as a hipGraph
  replay), forward and forward+backward.  Every view equals its single-view
  render bit for bit, one view equals the oracle bit for bit, and the batched
  backward equals the sum of the 32 single-view backward passes.
- Serving: the 8 resident scenes (10k to 5M Gaussians, 14.61M in all) under
  a Poisson stream of mixed resolutions with 20% forward+backward requests,
  including a 3840x2160 forward request, a forward+backward request on the
  5M scene and a 320x240 request.  Each result equals the same request served
  alone, bit for bit (images, depth, loss and gradients).
"""
import sys
import time

import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, look_at
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import SERVING_RESOLUTIONS, orbit_workload, serving_scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def orbit3m():
    say("[config4] generating 3M Gaussians and 32 orbit cameras")
    scene, cams = orbit_workload(3_000_000, 32, 6.0)
    return scene, scene.to_device(), cams


def test_orbit_3m_32_views_forward(orbit3m, rasterizer, oracle):
    import torch
    from lumen_amd.batch import MultiViewRenderer
    scene, dev, cams = orbit3m
    cfg = RenderConfig()
    mv = MultiViewRenderer(0)
    say("[config4] eager multi-view frame")
    eager = mv.render(dev, cams, cfg, want_stats=True)
    mv.synchronize()
    stats = eager.stats
    M = sum(s.total_splat_instances for s in stats)
    say(f"[config4] {len(cams)} views, {M} tile instances")
    assert M > 200_000_000
    ref_rgb, ref_depth = eager.rgb.clone(), eager.depth.clone()
    for rep in range(3):
        g = mv.render(dev, cams, cfg, graph=True)
        gst = mv.status(len(cams))
        torch.cuda.synchronize()
        assert [s.total_splat_instances for s in gst] == [s.total_splat_instances for s in stats]
        assert torch.equal(g.rgb, ref_rgb) and torch.equal(g.depth, ref_depth), rep
        say(f"[config4] graph replay {rep} equals the eager frame")
    for k, cam in enumerate(cams):
        one = rasterizer.render(dev, cam, cfg)
        torch.cuda.synchronize()
        assert one.visible_splats == stats[k].visible_splats, k
        assert one.total_splat_instances == stats[k].total_splat_instances, k
        assert torch.equal(ref_rgb[3 * k * 1920 * 1080:3 * (k + 1) * 1920 * 1080].view(1080, 1920, 3),
                           one.rgb), k
        assert torch.equal(ref_depth[k * 1920 * 1080:(k + 1) * 1920 * 1080].view(1080, 1920),
                           one.depth), k
    say("[config4] every view equals its single-view render")
    k = 11
    t0 = time.perf_counter()
    ref = oracle.rasterize(scene, cams[k], cfg)
    say(f"[config4] oracle view {k}: {time.perf_counter() - t0:.1f} s, M={ref['instances']}")
    assert ref["instances"] == stats[k].total_splat_instances
    view = ref_rgb[3 * k * 1920 * 1080:3 * (k + 1) * 1920 * 1080].view(1080, 1920, 3)
    np.testing.assert_array_equal(view.cpu().numpy(), ref["rgb"])
    np.testing.assert_array_equal(ref_depth[k * 1920 * 1080:(k + 1) * 1920 * 1080]
                                  .view(1080, 1920).cpu().numpy(), ref["depth"])
    mv.close()


def min_camera_depth(scene, cams):
    """Per Gaussian, its smallest camera-space depth over the views whose depth
    range holds it (float64, R(q)^T (p - eye))."""
    from scipy.spatial.transform import Rotation
    pos = scene.planes[0:3].T.astype(np.float64)
    out = np.full(scene.count, np.inf)
    for cam in cams:
        q = np.asarray(cam.rotation, np.float64)
        R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
        z = (pos - np.asarray(cam.position, np.float64)) @ R[:, 2]
        z[(z < cam.config.near_plane) | (z > cam.config.far_plane)] = np.inf
        out = np.minimum(out, z)
    return out


def test_orbit_3m_32_views_backward(orbit3m, rasterizer):
    """The batched backward of the 32-view frame against the 32 single-view
    backward passes summed in view order.  Each view's partial gradients are
    the same; the batch adds the views' contributions inside one thread per
    Gaussian (and forms the covariance adjoint once from the summed dL/dSigma),
    so the two agree to float rounding: relative L2 error below 1e-5 for every
    parameter group, over the Gaussians at least 0.2 in front of every camera
    that sees them.  A few Gaussians of this orbit sit about 1 cm in front of a
    camera (projective Jacobian ~1e5): their position, scale and rotation
    gradients are ill-conditioned in float32 (huge intermediate dL/dSigma
    cancels): any two summation orders of them differ at the level of their
    own size, so for them the check is that the two results are finite and
    agree within half of sum_v |g_v| (a missing or doubled view breaks that).
    Opacity and SH gradients, which do not pass through the projection, agree
    to 1e-5 everywhere."""
    import torch
    from lumen_amd.batch import MultiViewRenderer, rasterize_views
    from lumen_amd.training import GaussianGradients, rasterize
    scene, dev, cams = orbit3m
    n = scene.count
    cfg = RenderConfig()
    mv = MultiViewRenderer(0)
    planes = dev.planes.detach().clone().requires_grad_(True)
    sh = dev.sh.detach().clone().requires_grad_(True)
    vs = GaussianGradients(n)
    say("[config4] batched forward+backward")
    rgb, depth = rasterize_views(planes, sh, cams, cfg, mv, viewspace=vs)
    P = 1920 * 1080
    g = torch.Generator(device="cuda").manual_seed(4)
    w = torch.randn(rgb.numel(), generator=g, device="cuda") * 1e-3
    (rgb * w).sum().backward()
    torch.cuda.synchronize()
    gb = torch.cat([planes.grad.reshape(-1), sh.grad.reshape(-1)])
    mv.close()
    vs1 = GaussianGradients(n)
    S = torch.zeros_like(gb)
    A = torch.zeros_like(gb)
    say("[config4] 32 single-view forward+backward passes")
    for k, cam in enumerate(cams):
        planes.grad = None
        sh.grad = None
        r, _ = rasterize(planes, sh, cam, cfg, rasterizer=rasterizer, viewspace=vs1)
        (r.reshape(-1) * w[3 * k * P: 3 * (k + 1) * P]).sum().backward()
        gv = torch.cat([planes.grad.reshape(-1), sh.grad.reshape(-1)])
        S += gv
        A += gv.abs()
    torch.cuda.synchronize()
    assert torch.equal(vs.observation_count, vs1.observation_count)
    assert torch.allclose(vs.viewspace_grad_2d, vs1.viewspace_grad_2d, rtol=1e-5, atol=1e-6)
    assert int(vs.observation_count.max()) > 16
    near = torch.from_numpy(min_camera_depth(scene, cams) < 0.2).cuda()
    say(f"[config4] {int(near.sum())} Gaussians closer than 0.2 to a camera")
    groups = {"pos": (0, 3), "scale": (3, 6), "rot": (6, 10), "opacity": (10, 11)}
    d = gb.double() - S.double()
    for name, (a, b) in groups.items():
        sl = slice(a * n, b * n)
        keep = ~near.repeat(b - a) if name != "opacity" else torch.ones_like(near)
        err = float(d[sl][keep].norm() / S[sl][keep].double().norm())
        worst = float((d[sl].abs() / (A[sl].double() + 1e-30)).max())
        say(f"[config4] {name}: rel L2 err {err:.2e} (well-conditioned), "
            f"max |batched - summed| / sum|g_v| {worst:.2e} (all)")
        assert err < 1e-5, name
        assert bool(torch.isfinite(gb[sl]).all()) and worst < 0.5, name
    err = float(d[11 * n:].norm() / S[11 * n:].double().norm())
    say(f"[config4] sh: rel L2 err {err:.2e}")
    assert err < 1e-5


def test_serving_8_resident_scenes_mixed_stream():
    import torch
    from lumen_amd.serve import Request, RenderServer, latency_summary, poisson_requests
    say("[config5] generating the 8 resident scenes (14.61M Gaussians)")
    scenes = {k: v.to_device() for k, v in serving_scenes().items()}
    names = list(scenes)

    def target(req):
        H, W = req.camera.config.height, req.camera.config.width
        gen = torch.Generator(device="cuda").manual_seed(1000 + req.rid)
        return torch.rand((H, W, 3), generator=gen, device="cuda")

    reqs = poisson_requests(names, SERVING_RESOLUTIONS, rate=300.0, count=24,
                            backward_fraction=0.2, seed=55, targets=target)
    # the cases the stream must hold, appended at its end
    t_end = reqs[-1].arrival
    extra = [
        (names[7], (3840, 2160, 2800.0), False, (0.7, 1.9, 5.6)),   # 4K forward, 5M scene
        (names[7], (1280, 720, 933.3), True, (-2.2, 1.1, 5.9)),     # fwd+bwd on the 5M scene
        (names[0], (320, 240, 233.3), False, (1.4, -0.6, 4.2)),     # small forward
        (names[3], (3840, 2160, 2800.0), False, (5.0, 2.0, -1.0)),  # 4K forward, 1M scene
    ]
    for j, (name, (w, h, f), bwd, eye) in enumerate(extra):
        cam = look_at(CameraConfig(w, h, f, f, w / 2.0, h / 2.0, 0.01, 1000.0), eye)
        r = Request(len(reqs), name, cam, bwd, None, t_end + 0.001 * (j + 1))
        if bwd:
            r.target = target(r)
        reqs.append(r)
    assert any(r.backward and r.scene == names[7] for r in reqs)
    assert any(r.camera.config.width == 3840 for r in reqs)
    assert any(r.camera.config.width == 320 for r in reqs)
    server = RenderServer(scenes, slots=3, max_views=8)
    say(f"[config5] serving {len(reqs)} requests")
    t0 = time.perf_counter()
    done = server.run(reqs, realtime=True)
    summary = latency_summary(done, time.perf_counter() - t0)
    say(f"[config5] {summary} frames={server.frames}")
    assert len(done) == len(reqs)
    alone = RenderServer(scenes, slots=1, max_views=1)
    for r in sorted(done, key=lambda r: r.rid):
        solo = alone.serve_one(Request(r.rid, r.scene, r.camera, r.backward, r.target))
        assert torch.equal(solo.rgb, r.rgb), r.rid
        assert torch.equal(solo.depth, r.depth), r.rid
        if r.backward:
            assert solo.loss == r.loss, r.rid
            assert torch.equal(solo.grad_planes, r.grad_planes), r.rid
            assert torch.equal(solo.grad_sh, r.grad_sh), r.rid
        solo.rgb = solo.depth = solo.grad_planes = solo.grad_sh = None
        r.rgb = r.depth = r.grad_planes = r.grad_sh = None
    say(f"[config5] all {len(done)} requests equal alone and in the mix")
    server.close()
    alone.close()
