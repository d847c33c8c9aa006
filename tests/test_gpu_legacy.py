"""lumen_rasterize in LUMEN_EX_RASTER_LEGACY mode: the reference C library's
own renderer (c_api/src/lumen_c.c:380-575) on the GPU, bit for bit.

Through the C ABI (lumen_cloud_add_gaussian, lumen_framebuffer_create,
lumen_rasterizer_create, lumen_ex_rasterizer_set_mode, lumen_rasterize):
- the fixtures the reference's C library wrote (tests/golden/
  legacy_c_renderer.npz): the C test's cube over two frames, the synthetic
  sphere, a 512-splat lattice with exact depth ties over a coloured
  framebuffer, the same lattice with tile size 7 and other config values
  over two frames, and a camera that sees nothing: 8-bit RGB, depth, alpha
  and stats identical;
- random scenes of 20,000 splats (4,000 at tile size 1) against the oracle's
  restatement (itself equal to the C library), with tile sizes 16, 5, 64 and 1, opacities above
  1 and an early-termination alpha of 2 (alpha passing 1, weights turning
  negative);
- the mode is per rasterizer and switches back to the splat pipeline.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from lumen_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "legacy_c_renderer.npz")
CASES = ("c_test_cube", "sphere128", "scene512", "scene512b", "away")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def load_case(name):
    z = np.load(GOLDEN)
    return {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}


class LegacyRun:
    """A cloud, a framebuffer with given starting planes and a rasterizer in
    legacy mode, all through the C ABI."""

    def __init__(self, aos, W, H, cam, cfg, init):
        L = self.L = _lib.lib()
        aos = np.ascontiguousarray(aos, np.float32)
        self.cloud = L.lumen_cloud_create(max(len(aos), 1))
        for row in aos:
            g = _lib.LumenGaussian.from_buffer_copy(row.tobytes())
            assert L.lumen_cloud_add_gaussian(self.cloud, ctypes.byref(g)) == 0
        self.W, self.H = W, H
        self.fb = L.lumen_framebuffer_create(W, H)
        P = W * H
        f = self.fb.contents
        self.rgb = np.ctypeslib.as_array(f.rgb_data, shape=(P * 3,)).reshape(H, W, 3)
        self.depth = np.ctypeslib.as_array(f.depth_data, shape=(P,)).reshape(H, W)
        self.alpha = np.ctypeslib.as_array(f.alpha_data, shape=(P,)).reshape(H, W)
        self.rgb[...], self.depth[...], self.alpha[...] = init
        c = _lib.LumenConfig()
        c.image_width, c.image_height, c.tile_size = W, H, cfg[0]
        c.near_plane, c.far_plane, c.early_termination_alpha, c.gaussian_radius_sigma = cfg[1:]
        self.r = L.lumen_rasterizer_create(ctypes.byref(c))
        assert L.lumen_ex_rasterizer_set_mode(self.r, _lib.LUMEN_EX_RASTER_LEGACY) == 0
        assert L.lumen_ex_rasterizer_get_mode(self.r) == _lib.LUMEN_EX_RASTER_LEGACY
        self.cam = _lib.LumenCamera()
        self.cam.position = _lib.LumenVec3(*cam[0:3])
        self.cam.target = _lib.LumenVec3(*cam[3:6])
        self.cam.up = _lib.LumenVec3(*cam[6:9])
        self.cam.fov_x_degrees, self.cam.fov_y_degrees = cam[9], cam[10]
        self.cam.near_plane, self.cam.far_plane = cam[11], cam[12]

    def frame(self):
        st = _lib.LumenStats()
        rc = self.L.lumen_rasterize(self.r, self.cloud, ctypes.byref(self.cam), self.fb,
                                    ctypes.byref(st))
        assert rc == 0, self.L.lumen_ex_last_error()
        return np.array([st.total_splats, st.visible_splats, st.culled_splats, st.total_tiles,
                         st.touched_tiles], np.uint64)

    def close(self):
        self.L.lumen_framebuffer_destroy(self.fb)
        self.L.lumen_rasterizer_destroy(self.r)
        self.L.lumen_cloud_destroy(self.cloud)


@pytest.mark.parametrize("name", CASES)
def test_legacy_mode_reproduces_the_reference_c_library(name):
    c = load_case(name)
    W, H = (int(x) for x in c["size"])
    cfg = (int(c["tile"][0]), *[float(x) for x in c["config"]])
    run = LegacyRun(c["aos"], W, H, [float(x) for x in c["camera"]], cfg,
                    (c["init_rgb"], c["init_depth"], c["init_alpha"]))
    try:
        for k in range(int(c["frames"][0])):
            say(f"[legacy] {name} frame {k}")
            st = run.frame()
            np.testing.assert_array_equal(run.rgb, c[f"rgb{k}"])
            np.testing.assert_array_equal(run.depth.view(np.uint32), c[f"depth{k}"].view(np.uint32))
            np.testing.assert_array_equal(run.alpha.view(np.uint32), c[f"alpha{k}"].view(np.uint32))
            np.testing.assert_array_equal(st, c[f"stats{k}"])
    finally:
        run.close()


def random_scene_aos(n, seed, op_max):
    rng = np.random.default_rng(seed)
    aos = np.zeros((n, 59), np.float32)
    aos[:, 0:3] = rng.normal(0, 1.0, (n, 3))
    aos[:, 2] = np.round(aos[:, 2] * 4) / 4  # exact depth ties
    aos[:, 3:6] = np.exp(rng.normal(np.log(0.03), 0.6, (n, 3)))
    aos[:, 6] = 1.0
    aos[:, 10] = rng.uniform(0.0, op_max, n)
    aos[:, 11:14] = rng.normal(0, 1.0, (n, 3))
    return aos


@pytest.mark.parametrize("seed,tile,eta,sigma,op_max,n,W,H", [
    (1, 16, 0.99, 3.0, 1.0, 20_000, 640, 480),
    (2, 5, 0.9, 2.0, 1.2, 20_000, 640, 480),
    (3, 64, 2.0, 3.5, 3.0, 20_000, 640, 480),
    (4, 1, 0.99, 3.0, 1.0, 4_000, 200, 150)])
def test_legacy_mode_equals_the_oracle_on_random_scenes(oracle, seed, tile, eta, sigma, op_max,
                                                        n, W, H):
    aos = random_scene_aos(n, seed, op_max)
    cam = [0.4, -0.3, 4.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 60.0, 50.0, 0.1, 100.0]
    cfg = (tile, 0.2, 30.0, eta, sigma)
    rng = np.random.default_rng(seed + 100)
    init = (rng.integers(0, 256, (H, W, 3), dtype=np.uint8),
            np.full((H, W), 1000.0, np.float32), rng.uniform(0, 0.5, (H, W)).astype(np.float32))
    say(f"[legacy] random {n}, {W}x{H}, tile {tile}, eta {eta}, sigma {sigma}")
    run = LegacyRun(aos, W, H, cam, cfg, init)
    rgb, depth, alpha = (x.copy() for x in init)
    camera = {"pos": cam[0:3], "target": cam[3:6], "up": cam[6:9], "fov_x": cam[9],
              "fov_y": cam[10], "near": cam[11], "far": cam[12]}
    config = {"near": cfg[1], "far": cfg[2], "eta": cfg[3], "sigma": cfg[4], "tile": tile}
    try:
        for k in range(2):
            st = run.frame()
            ref = oracle.legacy_rasterize(aos, camera, config, rgb, depth, alpha)
            np.testing.assert_array_equal(run.rgb, rgb)
            np.testing.assert_array_equal(run.depth.view(np.uint32), depth.view(np.uint32))
            np.testing.assert_array_equal(run.alpha.view(np.uint32), alpha.view(np.uint32))
            assert st.tolist() == [ref["total"], ref["visible"], ref["culled"],
                                   ref["total_tiles"], ref["touched"]]
        assert ref["visible"] > n // 2
        if eta > 1.0:
            assert float(alpha.max()) > 1.0
    finally:
        run.close()


def test_mode_is_per_rasterizer_and_switches_back():
    L = _lib.lib()
    c = load_case("c_test_cube")
    cam = [float(x) for x in c["camera"]]
    cfg = (16, 0.1, 50.0, 0.99, 3.0)
    init = (c["init_rgb"], c["init_depth"], c["init_alpha"])
    legacy = LegacyRun(c["aos"], 160, 120, cam, cfg, init)
    pipe = LegacyRun(c["aos"], 160, 120, cam, cfg, init)
    try:
        assert L.lumen_ex_rasterizer_set_mode(pipe.r, _lib.LUMEN_EX_RASTER_PIPELINE) == 0
        assert L.lumen_ex_rasterizer_get_mode(pipe.r) == _lib.LUMEN_EX_RASTER_PIPELINE
        assert L.lumen_ex_rasterizer_set_mode(pipe.r, 7) != 0
        legacy.frame()
        pipe.frame()
        np.testing.assert_array_equal(legacy.rgb, c["rgb0"])
        assert not np.array_equal(pipe.rgb, c["rgb0"])  # the splat pipeline's own pixels
        assert float(pipe.alpha[60, 80]) > 0.0
    finally:
        legacy.close()
        pipe.close()


def test_sdk_engine_in_legacy_mode_renders_the_c_library_frame():
    """LumenEngine(legacy=True) renders the C test's cube as the reference C
    library does (its fresh framebuffer is the fixture's first frame)."""
    from lumen_amd.sdk import Camera, GaussianCloud, LumenEngine, RasterizerConfig
    from lumen_amd.sdk.cloud import Gaussian3D
    c = load_case("c_test_cube")
    cloud = GaussianCloud([Gaussian3D(tuple(float(v) for v in g[0:3]), tuple(float(v) for v in g[3:6]),
                                      tuple(float(v) for v in g[6:10]), float(g[10]),
                                      [float(v) for v in g[11:59]]) for g in c["aos"]])
    cfg = RasterizerConfig(160, 120, 16, 0.1, 50.0, 0.99, 3.0)
    eng = LumenEngine(cfg, legacy=True)
    try:
        res = eng.render(cloud, Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 45.0,
                                       0.1, 50.0))
        assert res.rgb_bytes == c["rgb0"].tobytes()
        np.testing.assert_array_equal(np.array(res.alpha_values, np.float32).reshape(120, 160),
                                      c["alpha0"])
        assert res.stats["visible_splats"] == int(c["stats0"][1])
        assert res.stats["touched_tiles"] == int(c["stats0"][4])
        two = eng.render_batch(cloud, [Camera((0.0, 0.0, 3.0)), Camera((0.0, 0.0, 3.0))])
        assert two[0].rgb_bytes == two[1].rgb_bytes == res.rgb_bytes
    finally:
        eng.close()
