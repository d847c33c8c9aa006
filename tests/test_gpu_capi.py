"""The reference's 22-symbol C ABI on the GPU: lumen_rasterize renders
through the splat pipeline and composites into the caller's framebuffer.

test_c_api.c:127-172 (the reference's own test of lumen_rasterize) is
restated here, then the framebuffer is pinned to the oracle's render of the
same cloud and camera: 8-bit colour trunc(255 min(1, rgb)), alpha = 1 - T,
depth = D / alpha (composited over an empty framebuffer), and a second call
composites behind the first.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

from lumen_amd import _lib
from lumen_amd.camera import CameraConfig, PinholeCamera
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import GaussianScene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def config(L, w=160, h=120):
    c = L.lumen_default_config()
    c.image_width, c.image_height, c.tile_size = w, h, 16
    c.near_plane, c.far_plane = 0.1, 50.0
    c.early_termination_alpha, c.gaussian_radius_sigma = 0.99, 3.0
    return c


def camera(eye=(0.0, 0.0, 3.0)):
    cam = _lib.LumenCamera()
    cam.position = _lib.LumenVec3(*eye)
    cam.target = _lib.LumenVec3(0.0, 0.0, 0.0)
    cam.up = _lib.LumenVec3(0.0, 1.0, 0.0)
    cam.fov_x_degrees, cam.fov_y_degrees = 60.0, 45.0
    cam.near_plane, cam.far_plane = 0.1, 50.0
    return cam


_libm = ctypes.CDLL("libm.so.6")
_libm.tanf.restype = ctypes.c_float
_libm.tanf.argtypes = [ctypes.c_float]


def legacy_focal(H, fov_y_deg):
    """H / (2 tanf(fov / 2)) in float, as lumen_rasterize computes it."""
    f32 = np.float32
    fov = f32(f32(f32(fov_y_deg) * f32(math.pi)) / f32(180.0))
    return float(f32(H) / (f32(2.0) * f32(_libm.tanf(float(f32(fov * f32(0.5)))))))


def legacy_look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """The camera-to-world quaternion lumen_rasterize derives (columns right,
    down, forward; Shepperd's method in double, then float)."""
    f = [float(np.float32(t)) - float(np.float32(e)) for t, e in zip(target, eye)]
    fl = math.sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2])
    f = [x / fl for x in f]
    u = [float(np.float32(x)) for x in up]
    r = [f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0]]
    rl = math.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
    r = [x / rl for x in r]
    d = [f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]]
    m00, m01, m02 = r[0], d[0], f[0]
    m10, m11, m12 = r[1], d[1], f[1]
    m20, m21, m22 = r[2], d[2], f[2]
    tr = m00 + m11 + m22
    if tr > 0.0:
        s = 2.0 * math.sqrt(tr + 1.0)
        q = (0.25 * s, (m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s)
    elif m00 > m11 and m00 > m22:
        s = 2.0 * math.sqrt(1.0 + m00 - m11 - m22)
        q = ((m21 - m12) / s, 0.25 * s, (m01 + m10) / s, (m02 + m20) / s)
    elif m11 > m22:
        s = 2.0 * math.sqrt(1.0 + m11 - m00 - m22)
        q = ((m02 - m20) / s, (m01 + m10) / s, 0.25 * s, (m12 + m21) / s)
    else:
        s = 2.0 * math.sqrt(1.0 + m22 - m00 - m11)
        q = ((m10 - m01) / s, (m02 + m20) / s, (m12 + m21) / s, 0.25 * s)
    return np.array(q, np.float32)


def cloud_scene(L, handle):
    n = L.lumen_cloud_get_count(handle)
    g = _lib.LumenGaussian()
    pos, scale, rot, op, sh = [], [], [], [], np.zeros((n, 48), np.float32)
    for i in range(n):
        assert L.lumen_cloud_get_gaussian(handle, i, ctypes.byref(g)) == 0
        pos.append((g.position.x, g.position.y, g.position.z))
        scale.append((g.scale.x, g.scale.y, g.scale.z))
        rot.append((g.rotation.w, g.rotation.x, g.rotation.y, g.rotation.z))
        op.append(g.opacity)
        sh[i] = np.ctypeslib.as_array(g.sh_coeffs)
    planes = np.concatenate([np.array(pos, np.float32).T, np.array(scale, np.float32).T,
                             np.array(rot, np.float32).T, np.array(op, np.float32)[None]])
    return GaussianScene(np.ascontiguousarray(planes), sh)


def test_rasterization_pipeline_kat():
    """test_c_api.c:127-172."""
    L = _lib.lib()
    cfg = config(L)
    r = L.lumen_rasterizer_create(ctypes.byref(cfg))
    cube = L.lumen_create_synthetic_cube(4, 1.0)
    cam = camera()
    fb = L.lumen_framebuffer_create(cfg.image_width, cfg.image_height)
    st = _lib.LumenStats()
    assert L.lumen_rasterize(r, cube, ctypes.byref(cam), fb, ctypes.byref(st)) == 0
    say(f"[capi] visible {st.visible_splats} touched {st.touched_tiles} {st.rasterize_time_ms:.3f} ms")
    assert st.total_splats == 64
    assert st.visible_splats > 0
    assert st.touched_tiles > 0
    assert st.rasterize_time_ms >= 0.0
    center = (cfg.image_height // 2) * cfg.image_width + cfg.image_width // 2
    assert fb.contents.alpha_data[center] > 0.0
    L.lumen_framebuffer_destroy(fb)
    L.lumen_cloud_destroy(cube)
    L.lumen_rasterizer_destroy(r)


@pytest.mark.parametrize("eye", [(0.0, 0.0, 3.0), (1.2, 0.8, 2.4)])
def test_framebuffer_matches_oracle(oracle, eye):
    L = _lib.lib()
    W, H = 160, 120
    cfg = config(L, W, H)
    r = L.lumen_rasterizer_create(ctypes.byref(cfg))
    cloud = L.lumen_create_synthetic_sphere(300, 1.0)
    cam = camera(eye)
    fb = L.lumen_framebuffer_create(W, H)
    st = _lib.LumenStats()
    assert L.lumen_rasterize(r, cloud, ctypes.byref(cam), fb, ctypes.byref(st)) == 0
    P = W * H
    rgb8 = np.ctypeslib.as_array(fb.contents.rgb_data, shape=(3 * P,)).reshape(H, W, 3).copy()
    alpha = np.ctypeslib.as_array(fb.contents.alpha_data, shape=(P,)).reshape(H, W).copy()
    depth = np.ctypeslib.as_array(fb.contents.depth_data, shape=(P,)).reshape(H, W).copy()
    # the oracle, with the camera lumen_rasterize derives (focal from the
    # vertical field of view, centre principal point, config clip planes)
    f = legacy_focal(H, 45.0)
    pc = PinholeCamera(CameraConfig(W, H, f, f, W / 2.0, H / 2.0, 0.1, 50.0),
                       np.array(eye, np.float32), legacy_look_at(eye))
    ref = oracle.rasterize(cloud_scene(L, cloud), pc, RenderConfig())
    assert st.visible_splats == ref["visible"]
    exp8 = (np.minimum(ref["rgb"], 1.0) * 255.0).astype(np.uint8)
    diff = np.abs(rgb8.astype(int) - exp8.astype(int))
    say(f"[capi] eye {eye}: max |rgb8 - oracle| = {diff.max()}, mismatched {(diff > 0).mean():.2e}")
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
    a_ref = 1.0 - ref["t_final"]
    assert np.max(np.abs(alpha - np.where(a_ref > 0, a_ref, 0.0))) <= 1e-4
    hit = a_ref > 1e-3
    np.testing.assert_allclose(depth[hit], ref["depth"][hit] / a_ref[hit], rtol=1e-3)
    assert np.all(depth[a_ref <= 0] == 1000.0)
    # a second frame composites behind the first
    a0 = alpha.copy()
    assert L.lumen_rasterize(r, cloud, ctypes.byref(cam), fb, ctypes.byref(st)) == 0
    alpha2 = np.ctypeslib.as_array(fb.contents.alpha_data, shape=(P,)).reshape(H, W)
    np.testing.assert_allclose(alpha2[hit], (a0 + (1 - a0) * a_ref)[hit], rtol=1e-5, atol=1e-6)
    L.lumen_framebuffer_destroy(fb)
    L.lumen_cloud_destroy(cloud)
    L.lumen_rasterizer_destroy(r)
