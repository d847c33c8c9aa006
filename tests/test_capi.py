"""liblumen.so C ABI: exports, layouts and the host-side behaviour of lumen.h.

Mirrors c_api/test/test_c_api.c:26-125 (version, config, cloud lifecycle,
synthetic scenes, framebuffer).  Where the reference's own C library has
been compiled into oracle/_ref, the synthetic clouds are compared with it
byte for byte.  None of these tests touches the GPU.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from lumen_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(name):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lumen_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    return _lib.lib()


def test_every_declared_symbol_is_exported(L):
    names = header_functions("lumen.h") + header_functions("lumen_ex.h")
    assert len(header_functions("lumen.h")) == 22
    for n in names:
        assert hasattr(L, n), n
    assert sorted(header_functions("lumen.h")) == sorted(_lib.LEGACY_SYMBOLS)


def test_struct_layouts():
    assert ctypes.sizeof(_lib.LumenConfig) == 28
    assert ctypes.sizeof(_lib.LumenVec3) == 12
    assert ctypes.sizeof(_lib.LumenQuat) == 16
    assert ctypes.sizeof(_lib.LumenCamera) == 52
    assert ctypes.sizeof(_lib.LumenGaussian) == 236
    assert ctypes.sizeof(_lib.LumenFrameBuffer) == 32
    assert ctypes.sizeof(_lib.LumenStats) == 48
    assert ctypes.sizeof(_lib.LumenExCamera) == 60
    assert ctypes.sizeof(_lib.LumenExRenderConfig) == 32


def test_version(L):
    assert L.lumen_get_version() == b"0.1.0"
    assert (L.lumen_get_version_major(), L.lumen_get_version_minor(),
            L.lumen_get_version_patch()) == (0, 1, 0)
    assert L.lumen_ex_abi_version() == 2


def test_config(L):
    c = L.lumen_default_config()
    assert (c.image_width, c.image_height, c.tile_size) == (800, 600, 16)
    assert 0 < c.near_plane < c.far_plane
    assert c.early_termination_alpha > 0.9


def _gaussian(x=1.0):
    g = _lib.LumenGaussian()
    g.position = _lib.LumenVec3(x, 2.0, 3.0)
    g.scale = _lib.LumenVec3(0.1, 0.1, 0.1)
    g.rotation = _lib.LumenQuat(1.0, 0.0, 0.0, 0.0)
    g.opacity = 0.9
    return g


def test_cloud_lifecycle(L):
    c = L.lumen_cloud_create(8)
    assert c and L.lumen_cloud_get_count(c) == 0 and L.lumen_cloud_get_capacity(c) >= 8
    assert L.lumen_cloud_add_gaussian(c, ctypes.byref(_gaussian())) == 0
    out = _lib.LumenGaussian()
    assert L.lumen_cloud_get_gaussian(c, 0, ctypes.byref(out)) == 0
    assert (out.position.x, out.position.y, out.position.z) == (1.0, 2.0, 3.0)
    for i in range(30):
        L.lumen_cloud_add_gaussian(c, ctypes.byref(_gaussian(float(i))))
    assert L.lumen_cloud_get_count(c) == 31 and L.lumen_cloud_get_capacity(c) >= 31
    assert L.lumen_cloud_get_capacity(c) == 32  # doubling from 8: 8 -> 16 -> 32
    assert L.lumen_cloud_get_gaussian(c, 31, ctypes.byref(out)) == -1
    L.lumen_cloud_clear(c)
    assert L.lumen_cloud_get_count(c) == 0
    L.lumen_cloud_destroy(c)


def test_null_handling(L):
    assert L.lumen_cloud_get_count(None) == 0 and L.lumen_cloud_get_capacity(None) == 0
    assert L.lumen_cloud_reserve(None, 4) == -1
    assert L.lumen_cloud_add_gaussian(None, None) == -1
    L.lumen_cloud_destroy(None)
    L.lumen_framebuffer_destroy(None)
    L.lumen_rasterizer_destroy(None)
    assert L.lumen_rasterizer_set_config(None, None) == -1
    assert L.lumen_rasterize(None, None, None, None, None) == -1
    assert not L.lumen_framebuffer_create(0, 10)
    c = L.lumen_cloud_create(0)
    assert L.lumen_cloud_get_capacity(c) == 16
    L.lumen_cloud_destroy(c)


def test_framebuffer(L):
    fb = L.lumen_framebuffer_create(128, 96)
    assert fb and fb.contents.width == 128 and fb.contents.height == 96
    assert fb.contents.depth_data[0] == 1000.0 and fb.contents.alpha_data[0] == 0.0
    L.lumen_framebuffer_clear(fb, 42, 84, 126)
    assert list(fb.contents.rgb_data[0:3]) == [42, 84, 126]
    assert fb.contents.rgb_data[3 * (128 * 96) - 1] == 126
    L.lumen_framebuffer_destroy(fb)


def test_rasterizer_handle_without_gpu(L):
    cfg = L.lumen_default_config()
    r = L.lumen_rasterizer_create(ctypes.byref(cfg))
    assert r
    cfg.tile_size = 8
    assert L.lumen_rasterizer_set_config(r, ctypes.byref(cfg)) == 0
    L.lumen_rasterizer_destroy(r)
    r = L.lumen_rasterizer_create(None)
    assert r
    L.lumen_rasterizer_destroy(r)


def _cloud_bytes(L, handle):
    n = L.lumen_cloud_get_count(handle)
    out = bytearray()
    g = _lib.LumenGaussian()
    for i in range(n):
        assert L.lumen_cloud_get_gaussian(handle, i, ctypes.byref(g)) == 0
        out += bytes(g)
    return bytes(out)


def _ref_lib():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.REF_PATH):
        pytest.skip("reference C library not built (make -C oracle ref)")
    R = ctypes.CDLL(pyoracle.REF_PATH)
    R.lumen_create_synthetic_cube.restype = ctypes.c_void_p
    R.lumen_create_synthetic_cube.argtypes = [ctypes.c_size_t, ctypes.c_float]
    R.lumen_create_synthetic_sphere.restype = ctypes.c_void_p
    R.lumen_create_synthetic_sphere.argtypes = [ctypes.c_size_t, ctypes.c_float]
    R.lumen_cloud_get_count.restype = ctypes.c_size_t
    R.lumen_cloud_get_count.argtypes = [ctypes.c_void_p]
    R.lumen_cloud_get_gaussian.restype = ctypes.c_int
    R.lumen_cloud_get_gaussian.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.POINTER(_lib.LumenGaussian)]
    R.lumen_cloud_destroy.argtypes = [ctypes.c_void_p]
    return R


@pytest.mark.parametrize("n,size", [(4, 2.0), (4, 1.0), (1, 1.0), (8, 0.3)])
def test_synthetic_cube_matches_reference_library(L, n, size):
    R = _ref_lib()
    a, b = L.lumen_create_synthetic_cube(n, size), R.lumen_create_synthetic_cube(n, size)
    assert L.lumen_cloud_get_count(a) == n ** 3
    assert _cloud_bytes(L, a) == _cloud_bytes(R, b)
    L.lumen_cloud_destroy(a)
    R.lumen_cloud_destroy(b)


@pytest.mark.parametrize("n,radius", [(128, 1.5), (1, 1.0), (1000, 3.0)])
def test_synthetic_sphere_matches_reference_library(L, n, radius):
    R = _ref_lib()
    a, b = L.lumen_create_synthetic_sphere(n, radius), R.lumen_create_synthetic_sphere(n, radius)
    assert L.lumen_cloud_get_count(a) == n
    A = np.frombuffer(_cloud_bytes(L, a), np.float32)
    B = np.frombuffer(_cloud_bytes(R, b), np.float32)
    # sinf/cosf/acosf come from the same libm; the expressions are the same
    np.testing.assert_array_equal(A, B)
    L.lumen_cloud_destroy(a)
    R.lumen_cloud_destroy(b)


def test_workspace_size_without_gpu():
    """lumen_ex_workspace_size is host arithmetic: it grows with the splats,
    instances, views and the backward, and rejects what lumen_ex_render would."""
    from lumen_amd._lib import LumenError
    from lumen_amd.camera import CameraConfig, look_at
    from lumen_amd.raster import RenderConfig, workspace_size
    cam = look_at(CameraConfig(1920, 1080, 1400.0, 1400.0, 960.0, 540.0), (0.0, 0.0, 5.0))
    base = workspace_size(1_000_000, cam, max_instances=6_000_000)
    # at least the ping-pong keys and values and the 48-byte records
    assert base >= 2 * 12 * 6_000_000 + 48 * 1_000_000
    assert workspace_size(2_000_000, cam, max_instances=6_000_000) > base
    assert workspace_size(1_000_000, cam, max_instances=12_000_000) > base
    assert workspace_size(1_000_000, [cam, cam], max_instances=6_000_000) > base
    bwd = workspace_size(1_000_000, cam, max_instances=6_000_000, backward=True)
    assert bwd >= base + 40 * 6_000_000
    assert workspace_size(1_000_000, cam, max_instances=6_000_000, graph=True) > base
    assert workspace_size(0, cam) > 0
    with pytest.raises(LumenError):
        workspace_size(10, cam, RenderConfig(tile_width=10, tile_height=10))
    with pytest.raises(LumenError):
        workspace_size(10, [])
