"""The Python SDK surface (lumen_amd.sdk) and scene I/O (lumen_amd.ply) on the
CPU: python/tests/test_client.py's camera, cloud and PLY cases restated, the
Mojo core's PLY round trip (tests/test_ply.mojo:13-45), the parse_float
semantics (src/core/types.mojo:66-117) against a scalar restatement, and the
binary loader.  GPU rendering through the SDK is in test_gpu_sdk.py."""
import math
import os

import numpy as np
import pytest

from lumen_amd import ply
from lumen_amd.camera import quaternion_to_rotation
from lumen_amd.scene import GaussianScene, random_scene
from lumen_amd.sdk import (Camera, CameraTrajectory, Gaussian3D, GaussianCloud, ascii_render,
                           export_depth_visualization, export_ppm)


# ---- python/tests/test_client.py:23-80 ---------------------------------------

def test_camera_forward():
    f = Camera(position=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0)).forward_vector()
    assert f == pytest.approx((0.0, 0.0, -1.0), abs=1e-5)


def test_view_matrix_identity():
    mat = Camera(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, -1.0)).view_matrix_4x4()
    assert len(mat) == 4 and len(mat[0]) == 4
    assert mat[3][3] == pytest.approx(1.0)
    # looking down -Z with +Y up: the identity rotation
    np.testing.assert_allclose(np.array(mat)[:3, :3], np.eye(3), atol=1e-12)


def test_orbit_trajectory():
    cams = CameraTrajectory.orbit(target=(0.0, 0.0, 0.0), radius=4.0, num_frames=12)
    assert len(cams) == 12
    for c in cams:
        assert math.hypot(c.position[0], c.position[2]) > 0.5
        # on the sphere of the radius, at 15 degrees elevation
        assert math.sqrt(sum(v * v for v in c.position)) == pytest.approx(4.0)
        assert c.position[1] == pytest.approx(4.0 * math.sin(math.radians(15.0)))
    assert cams[0].position[0] == pytest.approx(0.0, abs=1e-12) and cams[0].position[2] > 0


def test_camera_interpolation():
    mid = CameraTrajectory.interpolate(Camera(position=(0.0, 0.0, 2.0)),
                                       Camera(position=(0.0, 0.0, 4.0)), 0.5)
    assert mid.position[2] == pytest.approx(3.0)
    assert CameraTrajectory.interpolate(Camera(position=(0.0, 0.0, 2.0)),
                                        Camera(position=(0.0, 0.0, 4.0)), 7.0).position[2] == 4.0


def test_cube_generation():
    cube = GaussianCloud.synthetic_cube(points_per_axis=3, cube_size=2.0)
    assert len(cube) == 27
    lo, hi = cube.bounding_box()
    assert lo[0] == pytest.approx(-1.0) and hi[0] == pytest.approx(1.0)
    assert cube.centroid() == pytest.approx((0.0, 0.0, 0.0), abs=1e-12)
    # colour of the far corner through the DC term: 0.5 + C0 * sh = 1
    assert 0.5 + 0.2820947917 * cube.splats[-1].sh_coeffs[0] == pytest.approx(1.0)


def test_sphere_generation():
    sphere = GaussianCloud.synthetic_sphere(point_count=50, radius=2.0)
    assert len(sphere) == 50
    for g in sphere.splats:
        assert math.sqrt(sum(v * v for v in g.position)) == pytest.approx(2.0)


def test_ply_roundtrip(tmp_path):
    cloud = GaussianCloud.synthetic_cube(points_per_axis=2, cube_size=1.0)
    path = str(tmp_path / "c.ply")
    cloud.save_ply(path)
    loaded = GaussianCloud.load_ply(path)
    assert len(loaded) == len(cloud)
    for a, b in zip(loaded.splats, cloud.splats):
        assert a.position == pytest.approx(b.position, abs=1e-6)
        assert a.sh_coeffs[:3] == pytest.approx(b.sh_coeffs[:3], abs=1e-6)
        assert a.opacity == pytest.approx(b.opacity, abs=1e-6)


def test_visualizer(tmp_path):
    w, h = 40, 30
    rgb = bytes(range(256)) * ((3 * w * h) // 256) + bytes((3 * w * h) % 256)
    p = str(tmp_path / "a.ppm")
    export_ppm(p, w, h, rgb)
    data = open(p, "rb").read()
    assert data.startswith(b"P6\n40 30\n255\n") and data.endswith(rgb)
    depth = [1000.0] * (w * h)
    depth[0], depth[1], depth[2] = 0.5, 10.0, 5.25
    d = str(tmp_path / "d.ppm")
    export_depth_visualization(d, w, h, depth)
    px = np.frombuffer(open(d, "rb").read()[len(b"P6\n40 30\n255\n"):], np.uint8).reshape(-1, 3)
    # near -> intensity 255; far and empty -> 0; halfway -> int(127.5) = 127
    assert px[0].tolist() == [255, 204, int(255 - 255 * 0.7)]
    assert px[1].tolist() == [0, 0, 255] and px[3].tolist() == [0, 0, 255]
    assert px[2].tolist() == [127, int(127 * 0.8), int(255 - 127 * 0.7)]
    alpha = [0.0] * (w * h)
    alpha[14 * w + 20] = 1.0  # sampled by row 7 (y = 7 * 30/15), column 10
    art = ascii_render(w, h, alpha, 20, 15)
    lines = art.split("\n")
    assert len(lines) == 15 and all(len(x) == 20 for x in lines)
    assert lines[7][10] == "@" and art.count("@") == 1 and set(art) <= set(" @\n")


# ---- the GPU path's camera -------------------------------------------------

def test_rotation_quaternion_faces_the_target():
    q = Camera(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, 1.0), up=(0.0, -1.0, 0.0))
    np.testing.assert_array_equal(q.rotation_quaternion(), np.array([1, 0, 0, 0], np.float32))
    for eye in [(0.0, 0.0, 3.0), (1.2, 0.8, 2.4), (-3.0, 1.0, -0.5), (0.0, 5.0, 0.01)]:
        cam = Camera(position=eye, target=(0.1, -0.2, 0.0))
        R = quaternion_to_rotation(cam.rotation_quaternion())
        f = np.array(cam.forward_vector())
        np.testing.assert_allclose(R @ [0.0, 0.0, 1.0], f, atol=1e-6)   # +Z forward
        right = R @ [1.0, 0.0, 0.0]
        assert abs(right[1]) < 1e-6                                      # level horizon
        assert (R @ [0.0, 1.0, 0.0])[1] < 0.0 or abs(f[1]) > 0.99        # +Y down


def test_focal_length_from_fov():
    f = Camera(fov_y_deg=45.0).focal_length(600)
    assert f == pytest.approx(600.0 / (2.0 * math.tan(math.radians(22.5))), rel=1e-6)
    pc = Camera(position=(0.0, 0.0, 3.0)).to_pinhole(80, 60, 0.1, 50.0)
    assert (pc.config.principal_x, pc.config.principal_y) == (40.0, 30.0)
    assert (pc.config.near_plane, pc.config.far_plane) == (0.1, 50.0)


def test_cloud_layouts():
    cloud = GaussianCloud.synthetic_sphere(10, 1.5)
    a = cloud.aos()
    assert a.shape == (10, 59) and a.dtype == np.float32
    sc = cloud.to_scene()
    back = GaussianCloud.from_scene(sc)
    np.testing.assert_array_equal(back.aos(), a)
    np.testing.assert_array_equal(sc.planes[10], np.float32(0.9))


# ---- scene I/O: src/scene/ply.mojo ----------------------------------------

def _mojo_parse_float(s: bytes) -> np.float32:
    """parse_float (src/core/types.mojo:66-117), one character at a time."""
    f32 = np.float32
    sign, val, div = f32(1), f32(0), f32(1)
    in_dec = in_exp = exp_first = False
    exp_sign, exp_val = f32(1), f32(0)
    for j, c in enumerate(s):
        if j == 0 and c in (45, 43):
            if c == 45:
                sign = f32(-1)
            continue
        if c in (101, 69):
            in_exp, exp_first = True, True
            continue
        if in_exp:
            if exp_first:
                exp_first = False
                if c == 45:
                    exp_sign = f32(-1)
                    continue
                if c == 43:
                    continue
            if 48 <= c <= 57:
                exp_val = f32(exp_val * f32(10) + f32(c - 48))
        elif c == 46:
            in_dec = True
        elif 48 <= c <= 57:
            if in_dec:
                div = f32(div * f32(10))
                val = f32(val + f32(c - 48) / div)
            else:
                val = f32(val * f32(10) + f32(c - 48))
    total = f32(sign * val)
    if in_exp:
        total = f32(total * np.power(f32(10), f32(exp_sign * exp_val), dtype=f32))
    return total


def test_parse_float_matches_the_scalar_loop():
    rng = np.random.default_rng(5)
    toks = [b"0.95", b"-4.0", b"1e-05", b"+2.5E2", b"7", b"-0.000123", b"", b"-", b"3.",
            b".5", b"1.5e+3", b"12345678.9", b"nan"]
    for v in rng.normal(0, 10, 300):
        toks.append(("%.9g" % v).encode())
        toks.append(("%.6f" % (v / 1000)).encode())
    got = ply.parse_float_tokens(np.array(toks, dtype=np.bytes_))
    want = np.array([_mojo_parse_float(t) for t in toks], np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # and it is within a few ulp of a correctly rounded parse for plain decimals
    plain = [t for t in toks[13:] if b"e" not in t]
    exact = np.array([float(t) for t in plain], np.float32)
    par = ply.parse_float_tokens(np.array(plain, dtype=np.bytes_))
    np.testing.assert_allclose(par, exact, rtol=1e-5, atol=1e-7)


def test_mojo_ply_roundtrip(tmp_path):
    """tests/test_ply.mojo:13-45."""
    sh = np.zeros((2, 48), np.float32)
    sh[:, 0:3] = (0.5, 0.2, 0.8)
    scene = GaussianScene.from_arrays([(1.0, 2.0, 3.0), (-4.0, 5.0, -6.0)],
                                      [(0.1, 0.2, 0.3), (0.4, 0.5, 0.6)],
                                      [(1.0, 0.0, 0.0, 0.0), (0.0, 1.0, 0.0, 0.0)],
                                      [0.95, 0.5], sh)
    path = str(tmp_path / "rt.ply")
    ply.write_ply_ascii(scene, path)
    back = ply.read_ply_ascii(path)
    assert back.count == 2
    np.testing.assert_allclose(back.planes[0:3, 0], (1.0, 2.0, 3.0), atol=1e-4)
    assert abs(back.planes[10, 0] - 0.95) < 1e-4
    np.testing.assert_allclose(back.planes[0:3, 1], (-4.0, 5.0, -6.0), atol=1e-4)
    np.testing.assert_allclose(back.planes, scene.planes, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(back.sh, scene.sh, rtol=1e-6, atol=1e-7)
    # the Mojo column order in the header and on the lines
    text = open(path).read().split("end_header\n")
    props = [l.split()[-1] for l in text[0].splitlines() if l.startswith("property")]
    assert tuple(props) == ply.MOJO_COLUMNS
    first = [float(x) for x in text[1].splitlines()[0].split()]
    assert first[3:6] == pytest.approx([0.5, 0.2, 0.8]) and first[6] == pytest.approx(0.95)


def test_read_ply_by_name_reads_the_sdk_order(tmp_path):
    cloud = GaussianCloud.synthetic_cube(3, 1.0)
    path = str(tmp_path / "sdk.ply")
    cloud.save_ply(path)
    sc = ply.read_ply(path)
    ref = GaussianCloud.load_ply(path).to_scene()
    np.testing.assert_array_equal(sc.planes, ref.planes)
    np.testing.assert_array_equal(sc.sh, ref.sh)


def test_binary_ply_roundtrip_is_exact(tmp_path):
    scene = random_scene(5000, 1.0, 77)
    path = str(tmp_path / "b.ply")
    ply.write_ply_binary(scene, path)
    back = ply.read_ply(path)
    np.testing.assert_array_equal(back.planes, scene.planes)
    np.testing.assert_array_equal(back.sh, scene.sh)
    assert 0 < os.path.getsize(path) - 5000 * 59 * 4 < 4096  # 236-byte records after the header


def test_binary_ply_3dgs_activations(tmp_path):
    scene = random_scene(100, 1.0, 78)
    raw = GaussianScene(scene.planes.copy(), scene.sh.copy())
    raw.planes[10] = np.log(scene.planes[10] / (1.0 - scene.planes[10]))
    raw.planes[3:6] = np.log(scene.planes[3:6])
    path = str(tmp_path / "g.ply")
    ply.write_ply_binary(raw, path)
    back = ply.read_ply(path, activation="3dgs")
    np.testing.assert_allclose(back.planes[10], scene.planes[10], rtol=1e-5)
    np.testing.assert_allclose(back.planes[3:6], scene.planes[3:6], rtol=1e-5)
    # f_rest is channel-major: coefficient k of channel c is f_rest[c*15 + k-1]
    with open(path, "rb") as f:
        fmt, elements, off = ply._read_header(f)
    names = [p[0] for p in elements[0][2]]
    rec = np.fromfile(path, dtype=np.dtype([(n, "<f4") for n in names]), offset=off)
    assert rec["f_rest_16"][3] == scene.sh[3, 3 * 2 + 1]


def test_read_ply_rejects_what_it_cannot_read(tmp_path):
    p = tmp_path / "bad.ply"
    p.write_text("ply\nformat binary_big_endian 1.0\nelement vertex 0\nproperty float x\nend_header\n")
    with pytest.raises(ValueError):
        ply.read_ply(str(p))
    p.write_text("not a ply\n")
    with pytest.raises(ValueError):
        ply.read_ply(str(p))
