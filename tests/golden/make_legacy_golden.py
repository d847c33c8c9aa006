"""Writes tests/golden/legacy_c_renderer.npz from the reference's own C library.

    python tests/golden/make_legacy_golden.py

Runs lumen_rasterize of oracle/_ref/liblumen_ref.so -- the reference's
c_api/src/lumen_c.c compiled from its source in this container by
`make -C oracle ref` -- on a few scenes and stores, per case, the inputs (the
AoS Gaussians, the camera, the config, the framebuffer's starting planes) and
what the C library leaves in the framebuffer after each frame (8-bit RGB,
depth, alpha) with its stats.  The GPU's legacy mode must reproduce these
bytes exactly (tests/test_gpu_legacy.py); the oracle's restatement of the
renderer must too (tests/test_legacy_oracle.py).

Cases:
  c_test_cube  the C test's scene: the 64-splat synthetic cube, 160x120,
               camera at (0, 0, 3) (c_api/test/test_c_api.c:127-172), two
               frames (the second accumulates onto the first)
  sphere128    the 128-splat synthetic sphere at 200x150
  scene512     512 splats on an 8^3 lattice with random scales, opacities
               and DC colours, the z coordinates shared by each layer (64
               splats per exact depth tie), viewed along the z axis at
               320x240 over a framebuffer cleared to (42, 84, 126)
  scene512b    the same lattice with z jitter from an oblique camera, tile
               size 7, sigma 2.5, early termination 0.95, near 0.5, far 40,
               256x192, two frames
  away         the cube seen from a camera looking away: nothing drawn
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lumen_amd import _lib  # noqa: E402  (the ABI's struct layouts only)

REF = os.path.join(ROOT, "oracle", "_ref", "liblumen_ref.so")
OUT = os.path.join(HERE, "legacy_c_renderer.npz")


def ref_lib():
    L = ctypes.CDLL(REF)
    L.lumen_cloud_create.restype = ctypes.c_void_p
    L.lumen_cloud_create.argtypes = [ctypes.c_size_t]
    L.lumen_cloud_destroy.argtypes = [ctypes.c_void_p]
    L.lumen_cloud_add_gaussian.argtypes = [ctypes.c_void_p, ctypes.POINTER(_lib.LumenGaussian)]
    L.lumen_cloud_get_count.restype = ctypes.c_size_t
    L.lumen_cloud_get_count.argtypes = [ctypes.c_void_p]
    L.lumen_cloud_get_gaussian.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.POINTER(_lib.LumenGaussian)]
    L.lumen_create_synthetic_cube.restype = ctypes.c_void_p
    L.lumen_create_synthetic_cube.argtypes = [ctypes.c_size_t, ctypes.c_float]
    L.lumen_create_synthetic_sphere.restype = ctypes.c_void_p
    L.lumen_create_synthetic_sphere.argtypes = [ctypes.c_size_t, ctypes.c_float]
    L.lumen_framebuffer_create.restype = ctypes.POINTER(_lib.LumenFrameBuffer)
    L.lumen_framebuffer_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.lumen_framebuffer_destroy.argtypes = [ctypes.POINTER(_lib.LumenFrameBuffer)]
    L.lumen_framebuffer_clear.argtypes = [ctypes.POINTER(_lib.LumenFrameBuffer), ctypes.c_uint8,
                                          ctypes.c_uint8, ctypes.c_uint8]
    L.lumen_rasterizer_create.restype = ctypes.c_void_p
    L.lumen_rasterizer_create.argtypes = [ctypes.POINTER(_lib.LumenConfig)]
    L.lumen_rasterizer_destroy.argtypes = [ctypes.c_void_p]
    L.lumen_rasterize.restype = ctypes.c_int
    L.lumen_rasterize.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_lib.LumenCamera),
                                  ctypes.POINTER(_lib.LumenFrameBuffer),
                                  ctypes.POINTER(_lib.LumenStats)]
    return L


def cloud_aos(L, cloud) -> np.ndarray:
    n = L.lumen_cloud_get_count(cloud)
    out = np.zeros((n, 59), np.float32)
    g = _lib.LumenGaussian()
    for i in range(n):
        assert L.lumen_cloud_get_gaussian(cloud, i, ctypes.byref(g)) == 0
        out[i] = np.frombuffer(bytes(g), np.float32)
    return out


def cloud_from(L, aos):
    c = L.lumen_cloud_create(max(len(aos), 1))
    for row in np.ascontiguousarray(aos, np.float32):
        g = _lib.LumenGaussian.from_buffer_copy(row.tobytes())
        assert L.lumen_cloud_add_gaussian(c, ctypes.byref(g)) == 0
    return c


def lattice512(seed, z_jitter):
    rng = np.random.default_rng(seed)
    ax = np.array([-1 + 2 * i / 7 for i in range(8)], np.float32)
    pos = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    pos[:, :2] += rng.uniform(-0.03, 0.03, (512, 2)).astype(np.float32)
    if z_jitter:
        pos[:, 2] += rng.uniform(-0.03, 0.03, 512).astype(np.float32)
    aos = np.zeros((512, 59), np.float32)
    aos[:, 0:3] = pos
    aos[:, 3:6] = rng.uniform(0.02, 0.12, (512, 3))
    q = rng.standard_normal((512, 4))
    aos[:, 6:10] = q / np.linalg.norm(q, axis=1, keepdims=True)
    aos[:, 10] = rng.uniform(0.05, 1.0, 512)
    aos[:, 11:59] = rng.normal(0.0, 0.1, (512, 48))
    aos[:, 11:14] = rng.normal(0.0, 0.8, (512, 3))
    return aos


def run_case(L, name, aos, W, H, cam, cfg, frames=1, clear=None):
    """cam: (pos, target, up, fov_x, fov_y, near, far); cfg: (tile, near, far, eta, sigma)."""
    cloud = cloud_from(L, aos)
    lc = _lib.LumenConfig()
    lc.image_width, lc.image_height, lc.tile_size = W, H, cfg[0]
    lc.near_plane, lc.far_plane, lc.early_termination_alpha, lc.gaussian_radius_sigma = cfg[1:]
    r = L.lumen_rasterizer_create(ctypes.byref(lc))
    c = _lib.LumenCamera()
    c.position, c.target, c.up = (_lib.LumenVec3(*cam[0]), _lib.LumenVec3(*cam[1]),
                                  _lib.LumenVec3(*cam[2]))
    c.fov_x_degrees, c.fov_y_degrees, c.near_plane, c.far_plane = cam[3:]
    fb = L.lumen_framebuffer_create(W, H)
    if clear is not None:
        L.lumen_framebuffer_clear(fb, *clear)
    P = W * H

    def planes():
        f = fb.contents
        return (np.ctypeslib.as_array(f.rgb_data, shape=(P * 3,)).reshape(H, W, 3).copy(),
                np.ctypeslib.as_array(f.depth_data, shape=(P,)).reshape(H, W).copy(),
                np.ctypeslib.as_array(f.alpha_data, shape=(P,)).reshape(H, W).copy())

    out = {"aos": np.ascontiguousarray(aos, np.float32), "size": np.array([W, H], np.uint32),
           "camera": np.array([*cam[0], *cam[1], *cam[2], *cam[3:]], np.float32),
           "config": np.array(cfg[1:], np.float32), "tile": np.array([cfg[0]], np.uint32)}
    out["init_rgb"], out["init_depth"], out["init_alpha"] = planes()
    for k in range(frames):
        st = _lib.LumenStats()
        assert L.lumen_rasterize(r, cloud, ctypes.byref(c), fb, ctypes.byref(st)) == 0
        out[f"rgb{k}"], out[f"depth{k}"], out[f"alpha{k}"] = planes()
        out[f"stats{k}"] = np.array([st.total_splats, st.visible_splats, st.culled_splats,
                                     st.total_tiles, st.touched_tiles], np.uint64)
        print(f"{name} frame {k}: stats {out[f'stats{k}'].tolist()}, "
              f"pixels drawn {int((out[f'alpha{k}'] > 0).sum())}")
    out["frames"] = np.array([frames], np.uint32)
    L.lumen_framebuffer_destroy(fb)
    L.lumen_rasterizer_destroy(r)
    L.lumen_cloud_destroy(cloud)
    return {f"{name}/{k}": v for k, v in out.items()}


def main():
    if not os.path.exists(REF):
        raise SystemExit(f"{REF} missing: run `make -C oracle ref` (needs the reference sources)")
    L = ref_lib()
    cube = L.lumen_create_synthetic_cube(4, 1.0)
    cube_aos = cloud_aos(L, cube)
    L.lumen_cloud_destroy(cube)
    sphere = L.lumen_create_synthetic_sphere(128, 1.0)
    sphere_aos = cloud_aos(L, sphere)
    L.lumen_cloud_destroy(sphere)
    data = {}
    up = (0.0, 1.0, 0.0)
    data.update(run_case(L, "c_test_cube", cube_aos, 160, 120,
                         ((0, 0, 3), (0, 0, 0), up, 60.0, 45.0, 0.1, 50.0),
                         (16, 0.1, 50.0, 0.99, 3.0), frames=2))
    data.update(run_case(L, "sphere128", sphere_aos, 200, 150,
                         ((0.5, 0.4, 3.5), (0, 0, 0), up, 60.0, 50.0, 0.1, 50.0),
                         (16, 0.1, 50.0, 0.99, 3.0)))
    data.update(run_case(L, "scene512", lattice512(71, False), 320, 240,
                         ((0, 0, 3.5), (0, 0, 0), up, 60.0, 50.0, 0.1, 100.0),
                         (16, 0.1, 100.0, 0.99, 3.0), clear=(42, 84, 126)))
    data.update(run_case(L, "scene512b", lattice512(72, True), 256, 192,
                         ((0.7, 0.5, 3.2), (0.1, -0.1, 0), up, 60.0, 55.0, 0.1, 100.0),
                         (7, 0.5, 40.0, 0.95, 2.5), frames=2))
    data.update(run_case(L, "away", cube_aos, 160, 120,
                         ((0, 0, 3), (0, 0, 6), up, 60.0, 45.0, 0.1, 50.0),
                         (16, 0.1, 50.0, 0.99, 3.0)))
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
