"""Pins the oracle's analytic backward against central finite differences of
its FP64 forward.  Every parameter of every Gaussian of a small scene is
checked; samples whose +-eps renders change any discrete decision (cull,
radius, tile rectangle, depth key, alpha skip/clamp/cut-off, termination,
colour clamp) are excluded and counted.
"""
import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, look_at
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import GaussianScene


def fd_scene(n=40, seed=3):
    rng = np.random.default_rng(seed)
    pos = rng.normal(0, 0.35, (n, 3))
    scale = np.exp(rng.normal(np.log(0.06), 0.3, (n, 3)))
    rot = rng.standard_normal((n, 4)) * rng.uniform(0.5, 2.0, (n, 1))
    op = rng.uniform(0.2, 0.9, n)
    sh = rng.normal(0, 0.08, (n, 48))
    sh[:, :3] = rng.normal(0.3, 0.3, (n, 3))
    return GaussianScene.from_arrays(pos, scale, rot, op, sh)


def fd_camera():
    return look_at(CameraConfig(48, 40, 60.0, 58.0, 24.0, 20.0, 0.01, 1000.0), (0.4, 0.3, 2.5))


def check_fd(oracle, scene, cam, cfg, seed=0, eps=1e-6):
    planes = scene.planes.astype(np.float64)
    sh = scene.sh.astype(np.float64)
    rng = np.random.default_rng(seed)
    H, W = cam.config.height, cam.config.width
    g_rgb = rng.normal(0, 1, (H, W, 3))
    g_d = rng.normal(0, 0.3, (H, W))
    rgb0, d0, sig0 = oracle.forward64(planes, sh, cam, cfg)
    grads = oracle.backward64(planes, sh, cam, cfg, g_rgb, g_d)

    def loss(p, s):
        rgb, d, sig = oracle.forward64(p, s, cam, cfg)
        return float(np.sum(rgb * g_rgb) + (np.sum(d * g_d) if cfg.compute_depth else 0.0)), sig

    checked = skipped = 0
    worst = 0.0
    scale_ref = max(np.max(np.abs(grads["planes"])), np.max(np.abs(grads["sh"])))
    for which, arr, garr in (("planes", planes, grads["planes"]), ("sh", sh, grads["sh"])):
        for idx in np.ndindex(arr.shape):
            h = eps * max(1.0, abs(arr[idx]))
            p1, p2 = arr.copy(), arr.copy()
            p1[idx] += h
            p2[idx] -= h
            if which == "planes":
                lp, sp = loss(p1, sh)
                lm, sm = loss(p2, sh)
            else:
                lp, sp = loss(planes, p1)
                lm, sm = loss(planes, p2)
            if sp != sig0 or sm != sig0:
                skipped += 1
                continue
            fd = (lp - lm) / (2 * h)
            an = garr[idx]
            err = abs(fd - an) / (1e-3 * scale_ref + abs(an))
            worst = max(worst, err)
            assert err < 1e-4, (which, idx, fd, an)
            checked += 1
    return checked, skipped, worst, grads


def test_backward64_matches_finite_differences(oracle):
    sc, cam = fd_scene(), fd_camera()
    checked, skipped, worst, g = check_fd(oracle, sc, cam, RenderConfig(tile_width=8, tile_height=8))
    assert checked > 0.95 * (checked + skipped)
    assert np.count_nonzero(g["planes"]) > 0.5 * g["planes"].size
    assert np.all(g["count"] <= 1)


def test_backward64_white_background_and_no_depth(oracle):
    sc, cam = fd_scene(24, 5), fd_camera()
    cfg = RenderConfig(white_background=True, compute_depth=False, max_sh_degree=2)
    checked, skipped, worst, g = check_fd(oracle, sc, cam, cfg, seed=1)
    assert checked > 0.95 * (checked + skipped)


def test_forward64_close_to_fp32_oracle(oracle):
    """The FP64 model is the FP32 path in exact arithmetic: same decisions,
    images within FP32 rounding."""
    sc, cam = fd_scene(200, 9), fd_camera()
    cfg = RenderConfig()
    rgb64, d64, _ = oracle.forward64(sc.planes.astype(np.float64), sc.sh.astype(np.float64), cam, cfg)
    r32 = oracle.rasterize(sc, cam, cfg)
    assert np.max(np.abs(rgb64 - r32["rgb"])) < 1e-4
    assert np.max(np.abs(d64 - r32["depth"])) < 1e-3


def test_viewspace_accumulation_semantics(oracle):
    """GaussianGradients.accumulate_viewspace_grad / get_average_viewspace_grad
    (src/optim/gradients.mojo:96-106; KAT tests/test_gradients.mojo:9-26):
    the sum over views divided by the number of views."""
    sc, cam = fd_scene(30, 2), fd_camera()
    cam2 = look_at(cam.config, (-0.4, 0.2, 2.6))
    cfg = RenderConfig()
    H, W = cam.config.height, cam.config.width
    g = np.ones((H, W, 3))
    gd = np.zeros((H, W))
    a = oracle.backward64(sc.planes, sc.sh, cam, cfg, g, gd)
    b = oracle.backward64(sc.planes, sc.sh, cam2, cfg, g, gd)
    vis = (a["count"] > 0) & (b["count"] > 0)
    assert vis.sum() > 10
    avg = (a["viewspace"] + b["viewspace"]) / (a["count"] + b["count"])
    assert np.all(avg[vis] == pytest.approx(((a["viewspace"] + b["viewspace"]) / 2)[vis]))
    # KAT: (0.04 + 0.06) / 2 = 0.05
    assert (0.04 + 0.06) / 2 == pytest.approx(0.05)
