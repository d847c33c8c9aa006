"""HIP path vs the CPU oracle, stage by stage, through the C ABI.

Bar: visible set, splat records, radii, unsorted and sorted keys/values and
tile ranges bit-exact; images and depth maps bit-exact as well (the compositor
evaluates the same FP32 expression trees and the same exp), which implies the
north-star tolerance of 1e-4 max-abs per channel, asserted separately.
"""
import math
import sys

import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, PinholeCamera, look_at
from lumen_amd.raster import RenderConfig, build_ranges
from lumen_amd.scene import GaussianScene, lattice_scene, random_scene, workload

pytestmark = pytest.mark.gpu

TOL = 1e-4  # north-star image/depth tolerance, max-abs per channel


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(rasterizer, oracle, scene, camera, config=None, exact_images=True):
    import torch
    config = config or RenderConfig()
    say(f"[parity] N={scene.count} {camera.config.width}x{camera.config.height} "
        f"tile={config.tile_width}x{config.tile_height} batch={config.batch_depth} "
        f"sh={config.max_sh_degree} white={config.white_background} depth={config.compute_depth}")
    ref = oracle.rasterize(scene, camera, config)
    out = rasterizer.render(scene, camera, config, keep_stages=True, save_pixels=True)
    torch.cuda.synchronize()
    V, M = ref["visible"], ref["instances"]
    assert out.visible_splats == V
    assert out.total_splat_instances == M
    if V:
        np.testing.assert_array_equal(rasterizer.stage("vis_to_orig"), ref["vis_to_orig"])
        np.testing.assert_array_equal(rasterizer.stage("radius"), ref["radius"])
        rec = rasterizer.stage("records")
        P = ref["proj"]
        # cx cy depth i00 i01 i11 r g b op
        got = np.stack([rec[:, 0], rec[:, 1], rec[:, 9], rec[:, 2], rec[:, 3] * np.float32(0.5),
                        rec[:, 4], rec[:, 6], rec[:, 7], rec[:, 8], rec[:, 5]], axis=1)
        np.testing.assert_array_equal(got, P)
    if M:
        np.testing.assert_array_equal(rasterizer.stage("keys_unsorted"), ref["keys_unsorted"])
        np.testing.assert_array_equal(rasterizer.stage("vals_unsorted"), ref["vals_unsorted"])
        np.testing.assert_array_equal(rasterizer.stage("keys"), ref["keys"])
        np.testing.assert_array_equal(rasterizer.stage("vals"), ref["vals"])
    np.testing.assert_array_equal(rasterizer.stage("ranges"), ref["ranges"])
    rgb = out.rgb.cpu().numpy()
    depth = out.depth.cpu().numpy()
    assert np.max(np.abs(rgb - ref["rgb"]), initial=0.0) <= TOL
    assert np.max(np.abs(depth - ref["depth"]), initial=0.0) <= TOL
    if exact_images:
        np.testing.assert_array_equal(rgb, ref["rgb"])
        np.testing.assert_array_equal(depth, ref["depth"])
        np.testing.assert_array_equal(rasterizer.stage("t_final").reshape(ref["t_final"].shape),
                                      ref["t_final"])
        np.testing.assert_array_equal(rasterizer.stage("n_eval").reshape(ref["n_eval"].shape),
                                      ref["n_eval"])
    # the same frame without the kept unsorted stages: single-view frames of
    # at least 2048 tiles then run the sort's first pass inside the key
    # emission (keys written in that pass's order), the others emit in the
    # reference's order; the sorted keys, values, ranges and pixels must not
    # change either way
    out2 = rasterizer.render(scene, camera, config, save_pixels=True)
    torch.cuda.synchronize()
    assert out.emit_passes == 0  # kept stages: the reference's emission order
    tiles = math.ceil(camera.config.width / config.tile_width) * \
        math.ceil(camera.config.height / config.tile_height)
    if tiles < 2048 or M == 0:
        assert out2.emit_passes == 0
    say(f"[parity]   unkept frame: {out2.sort_passes} passes, {out2.emit_passes} in emit")
    assert out2.sort_passes == out.sort_passes
    if M:
        np.testing.assert_array_equal(rasterizer.stage("keys"), ref["keys"])
        np.testing.assert_array_equal(rasterizer.stage("vals"), ref["vals"])
    np.testing.assert_array_equal(rasterizer.stage("ranges"), ref["ranges"])
    rgb2 = out2.rgb.cpu().numpy()
    if exact_images:
        np.testing.assert_array_equal(rgb2, ref["rgb"])
        np.testing.assert_array_equal(out2.depth.cpu().numpy(), ref["depth"])
        np.testing.assert_array_equal(rasterizer.stage("n_eval").reshape(ref["n_eval"].shape),
                                      ref["n_eval"])
    else:
        np.testing.assert_array_equal(rgb2, rgb)
    # and without the per-pixel state (the compositor instance that does not
    # track where pixels terminate): the same pixels
    out3 = rasterizer.render(scene, camera, config)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out3.rgb.cpu().numpy(), rgb2)
    np.testing.assert_array_equal(out3.depth.cpu().numpy(), out2.depth.cpu().numpy())
    return ref, out


def test_lattice_800x600(rasterizer, oracle):
    sc, cam = workload("lattice512_800x600")
    ref, _ = compare(rasterizer, oracle, sc, cam)
    assert ref["visible"] == 512


def test_random100k_1280x720(rasterizer, oracle):
    sc, cam = workload("random100k_1280x720")
    ref, out = compare(rasterizer, oracle, sc, cam)
    assert ref["instances"] > 400_000
    # 3600 tiles: the unkept frame's first pass runs in the key emission
    assert rasterizer.render(sc, cam, RenderConfig()).emit_passes == 1


def test_capacity_growth_and_reuse(oracle):
    """A fresh context renders a small cloud, a larger one and the small one
    again.  The plain frame enqueues plan and emission with the capacity its
    buffers already hold before it reads the totals back: the larger frame
    overflows that capacity (flagged on the device, both kernels run again
    after the buffers grow), the third fits inside it.  Every frame is the
    oracle's, stage by stage."""
    from lumen_amd.raster import Rasterizer
    cam = look_at(CameraConfig(640, 480, 600.0, 600.0, 320.0, 240.0, 0.01, 100.0), (0.0, 0.0, 5.0))
    small = random_scene(3_000, seed=2101)
    big = random_scene(40_000, seed=2102)
    R = Rasterizer(0)
    counts = []
    for sc in (small, big, small, big):
        ref, out = compare(R, oracle, sc, cam)
        counts.append(ref["instances"])
    assert counts[1] > 4 * counts[0] > 0


@pytest.mark.parametrize("tw,th,batch", [(8, 8, 64), (32, 8, 256), (16, 16, 64), (16, 16, 128),
                                         (16, 8, 128), (8, 16, 256)])
def test_tile_shapes_and_batches(rasterizer, oracle, tw, th, batch):
    sc = lattice_scene(1001)
    cam = look_at(CameraConfig(333, 217, 290.0, 290.0, 166.5, 108.5), (0.5, -0.3, 3.5))
    compare(rasterizer, oracle, sc, cam, RenderConfig(tile_width=tw, tile_height=th,
                                                      batch_depth=batch))


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_sh_degrees(rasterizer, oracle, deg):
    sc = random_scene(20_000, 1.0, 77)
    cam = look_at(CameraConfig(320, 240, 300.0, 300.0, 160.0, 120.0), (1.0, 2.0, 4.0))
    compare(rasterizer, oracle, sc, cam, RenderConfig(max_sh_degree=deg))


def test_white_background_no_depth(rasterizer, oracle):
    sc = lattice_scene(1001)
    cam = look_at(CameraConfig(300, 200, 260.0, 260.0, 150.0, 100.0), (0.0, 0.0, 4.0))
    compare(rasterizer, oracle, sc, cam, RenderConfig(white_background=True, compute_depth=False))


def test_single_splat_render(rasterizer, oracle):
    """tests/test_rasterizer.mojo:15-71 through the GPU path."""
    sh = np.zeros((1, 48), np.float32)
    sh[0, 0:3] = [1.77245, -1.77245, -1.77245]
    sc = GaussianScene.from_arrays([[0, 0, 5]], [[0.5] * 3], [[1, 0, 0, 0]], [1.0], sh)
    cam = PinholeCamera(CameraConfig(64, 64, 50.0, 50.0, 32.0, 32.0, 0.1, 100.0), [0, 0, 0],
                        [1, 0, 0, 0])
    _, out = compare(rasterizer, oracle, sc, cam)
    assert out.visible_splats == 1 and out.total_splat_instances > 0
    r, g, b = out.get_pixel_rgb(32, 32)
    assert r > 0.8 and g < 0.2 and b < 0.2
    assert out.get_pixel_rgb(0, 0) == (0.0, 0.0, 0.0)
    assert 4.5 < float(out.depth[32, 32]) < 5.5


def test_empty_offscreen_and_culled(rasterizer, oracle):
    """tests/test_regressions.mojo:115-146 plus a camera facing away."""
    cam = PinholeCamera(CameraConfig(32, 32, 20.0, 20.0, 16.0, 16.0, 0.1, 50.0), [0, 0, 0],
                        [1, 0, 0, 0])
    empty = GaussianScene(np.zeros((11, 0), np.float32), np.zeros((0, 48), np.float32))
    _, out = compare(rasterizer, oracle, empty, cam, RenderConfig(white_background=True))
    # pipeline.mojo:63-65 returns the zero-filled output before any background fill
    assert out.visible_splats == 0 and float(out.rgb.max()) == 0.0
    far = GaussianScene.from_arrays([[1e5, 1e5, 5]], [[0.1] * 3], [[1, 0, 0, 0]], [1.0],
                                    np.zeros((1, 48), np.float32))
    _, out = compare(rasterizer, oracle, far, cam)
    assert out.visible_splats == 0
    away = look_at(CameraConfig(64, 48, 60.0, 60.0, 32.0, 24.0), (0, 0, 4), (0, 0, 8))
    _, out = compare(rasterizer, oracle, lattice_scene(1001), away)
    assert out.visible_splats == 0


def test_camera_inside_cloud_and_large_splats(rasterizer, oracle):
    """Near-plane culling, splats covering the whole screen (long key runs)."""
    rng = np.random.default_rng(21)
    n = 3000
    pos = rng.normal(0, 1.0, (n, 3))
    scale = np.exp(rng.normal(math.log(0.05), 1.0, (n, 3)))
    scale[:5] = 3.0
    rot = rng.standard_normal((n, 4))
    op = 1 / (1 + np.exp(-rng.normal(0, 1.5, n)))
    sh = rng.normal(0, 0.3, (n, 48))
    sc = GaussianScene.from_arrays(pos, scale, rot, op, sh)
    cam = look_at(CameraConfig(400, 300, 300.0, 300.0, 200.0, 150.0, 0.05, 1000.0), (0.1, 0.2, 0.3),
                  (1.0, 0.5, 2.0))
    compare(rasterizer, oracle, sc, cam)


def test_thin_rotated_splats(rasterizer, oracle):
    """Needle-like splats (axis ratios 20:1 to 1000:1) at every angle, with
    opacities down to the 1/255 limit: their boxes meet strips their
    ellipses miss, and their conics straddle the condition-number bound up
    to which the compositors test the ellipse exactly.  Every stage and
    pixel stays the oracle's, in both tile shapes with two and one pixel per
    lane."""
    rng = np.random.default_rng(31)
    n = 4000
    pos = rng.uniform(-1.5, 1.5, (n, 3))
    pos[:, 2] *= 0.3
    long_ = np.exp(rng.uniform(math.log(0.05), math.log(1.2), n))
    ratio = np.exp(rng.uniform(math.log(20.0), math.log(1000.0), n))
    scale = np.stack([long_, long_ / ratio, long_ / ratio], axis=1)
    ang = rng.uniform(0, math.pi, n)  # rotation about the view axis (z)
    rot = np.stack([np.cos(ang / 2), np.zeros(n), np.zeros(n), np.sin(ang / 2)], axis=1)
    rot[n // 2:] = rng.standard_normal((n - n // 2, 4))  # and arbitrary 3D ones
    op = np.exp(rng.uniform(math.log(1.0 / 255.0), 0.0, n)).clip(0.0, 0.99)
    sh = rng.normal(0, 0.3, (n, 48))
    sc = GaussianScene.from_arrays(pos, scale, rot, op, sh)
    cam = look_at(CameraConfig(640, 480, 500.0, 500.0, 320.0, 240.0, 0.01, 100.0), (0.0, 0.0, 4.0))
    for tw, th in ((16, 16), (8, 8)):
        compare(rasterizer, oracle, sc, cam, RenderConfig(tile_width=tw, tile_height=th))


def test_unnormalized_and_degenerate_inputs(rasterizer, oracle):
    """Raw quaternions of any norm (incl. ~0 -> identity), scales below the
    1e-5 clamp, opacity at the 1/255 cull threshold."""
    rng = np.random.default_rng(8)
    n = 4000
    pos = rng.normal(0, 0.7, (n, 3))
    scale = np.exp(rng.normal(math.log(0.02), 1.0, (n, 3)))
    scale[:100, 1] = 0.0
    rot = rng.standard_normal((n, 4)) * rng.uniform(0.01, 10, (n, 1))
    rot[100:120] = 0.0
    op = rng.uniform(0, 1, n).astype(np.float32)
    op[200:220] = np.float32(0.00392156862)
    op[220:240] = np.nextafter(np.float32(0.00392156862), np.float32(0))
    sh = rng.normal(0, 0.5, (n, 48))
    sc = GaussianScene.from_arrays(pos, scale, rot, op, sh)
    cam = look_at(CameraConfig(256, 256, 250.0, 250.0, 128.0, 128.0), (0.0, 0.0, 3.0))
    compare(rasterizer, oracle, sc, cam)


def test_random1m_1920x1080_full_parity(rasterizer, oracle):
    """The 1M-splat 1080p workload, every stage bit-exact against the oracle."""
    sc, cam = workload("random1m_1920x1080")
    ref, out = compare(rasterizer, oracle, sc, cam)
    say(f"[parity] 1M: V={ref['visible']} M={ref['instances']} E={ref['evaluations']}")
    # 27-bit frame keys: the key emission runs the first of the 3 passes
    o = rasterizer.render(sc, cam, RenderConfig())
    assert (o.sort_passes, o.emit_passes) == (3, 1)


@pytest.mark.parametrize("eye", [(1.3, 3.9, 2.8), (1.3, 0.7, 4.8)])
def test_random1m_1920x1080_off_axis_camera(rasterizer, oracle, eye):
    """The 1M-splat 1080p scene from off-axis look-at cameras whose rotations
    are not their own inverses (2.26 and 3.00 rad): the camera transform,
    view rotation and every later stage bit-exact against the oracle (whose
    camera is pinned to the rotation library in tests/test_oracle_camera.py)."""
    sc, _ = workload("random1m_1920x1080")
    cfg = CameraConfig(1920, 1080, 1400.0, 1400.0, 960.0, 540.0, 0.01, 1000.0)
    ref, out = compare(rasterizer, oracle, sc, look_at(cfg, eye))
    assert ref["visible"] > 300_000
    say(f"[parity] 1M off-axis {eye}: V={ref['visible']} M={ref['instances']}")


# ---- sort and ranges through lumen_ex_radix_sort / lumen_ex_build_ranges --

def _lcg(n, seed=123456789):
    out = np.empty(n, np.uint64)
    s = seed
    for i in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out[i] = s
    return out


def _gpu_sort(rasterizer, keys):
    import torch
    k = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    v = torch.arange(keys.size, dtype=torch.int32).cuda()
    rasterizer.radix_sort(k, v)
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint64), v.cpu().numpy()


def test_radix_sort_lcg_kat(rasterizer, oracle):
    """tests/test_radix_sort.mojo:53-68: 1,000 full 64-bit LCG keys."""
    keys = _lcg(1000)
    k, v = _gpu_sort(rasterizer, keys)
    rk, rv = oracle.radix_sort(keys, np.arange(keys.size))
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(v, rv)


def test_radix_sort_composite_kat(rasterizer, oracle):
    """tests/test_radix_sort.mojo:14-50."""
    L = oracle.lib()
    keys = np.array([L.orc_encode_key(1, 5.0), L.orc_encode_key(0, 10.0),
                     L.orc_encode_key(0, 2.0), L.orc_encode_key(1, 1.0)], np.uint64)
    _, v = _gpu_sort(rasterizer, keys)
    assert v.tolist() == [2, 1, 3, 0]


# 2^25 + 1 keys: past the size where a sort block takes 16 keys per thread
# instead of 12 (block edges 3072 apart below it, 4096 above)
@pytest.mark.parametrize("n", [0, 1, 2, 3071, 3072, 3073, 4095, 4096, 4097, 3_000_001,
                               (1 << 25) + 1])
def test_radix_sort_matches_stable_argsort(rasterizer, n):
    rng = np.random.default_rng(n)
    hi = rng.integers(0, 1 << 13, n).astype(np.uint64) << np.uint64(32)
    lo = rng.integers(0, 20000, n).astype(np.uint64)  # many ties
    keys = hi | lo
    if n > 10:
        keys[: n // 3] = rng.integers(0, 2 ** 63, n // 3, dtype=np.int64).astype(np.uint64)
    say(f"[sort] n={n}")
    k, v = _gpu_sort(rasterizer, keys)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, order.astype(np.int32))


def test_radix_sort_past_2_pow_30_keys(rasterizer):
    """2^30 + 3 keys (8 digit passes of 64-bit keys with many ties), checked
    by their properties on the device: sorted, every output pair (key, value)
    is the input key at index value, the values are a permutation, and equal
    keys keep their input order."""
    import torch
    n = (1 << 30) + 3
    g = torch.Generator(device="cuda").manual_seed(30)
    say(f"[sort] n={n}: 64-bit look-back words")
    keys = (torch.randint(0, 1 << 16, (n,), generator=g, device="cuda") << 40) | \
        torch.randint(0, 1 << 10, (n,), generator=g, device="cuda")
    k = keys.clone()
    v = torch.arange(n, dtype=torch.int32, device="cuda")
    rasterizer.radix_sort(k, v)
    torch.cuda.synchronize()
    assert bool((k[1:] >= k[:-1]).all())
    ties = k[1:] == k[:-1]
    assert bool((v[1:][ties] > v[:-1][ties]).all())
    idx = v.long()
    assert bool((keys[idx] == k).all())
    del keys, k
    seen = torch.zeros(n, dtype=torch.uint8, device="cuda")
    seen[idx] = 1
    assert bool(seen.all()) and int(ties.sum()) > n // 2


def test_build_ranges_kat(rasterizer, oracle):
    """tests/test_binning.mojo:10-51."""
    import torch
    L = oracle.lib()
    keys = np.array([L.orc_encode_key(t, d) for t, d in
                     [(0, 1), (0, 2), (2, 3), (2, 4), (2, 5), (3, 6), (3, 7)]], np.uint64)
    r = build_ranges(torch.from_numpy(keys.view(np.int64)).cuda(), 4).cpu().numpy()
    assert r.tolist() == [[0, 2], [-1, -1], [2, 5], [5, 7]]


def test_wide_keys_take_the_64bit_sort(rasterizer, oracle):
    """Tile ids and depth keys that need more than 32 bits together: the
    frame is sorted as 64-bit keys in 8-bit digits (the other frames above as
    32-bit frame keys); both give the oracle's order."""
    sc = random_scene(30_000, 1000.0, 81)
    sc.planes[3:6] *= 1000.0
    cam = look_at(CameraConfig(640, 480, 600.0, 600.0, 320.0, 240.0, 0.01, 1.0e6),
                  (0.0, 0.0, 5000.0))
    ref, _ = compare(rasterizer, oracle, sc, cam)
    keys = ref["keys"]
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    depth = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    bits = int(tiles.max() - tiles.min()).bit_length() + int(depth.max() - depth.min()).bit_length()
    say(f"[parity] wide keys: {keys.size} keys, {bits} varying bits")
    assert bits > 32
