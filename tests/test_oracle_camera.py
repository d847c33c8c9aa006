"""Pins for the oracle's camera transform and tile rectangles.

The configs' cameras are look_at views from (0, 0, z): q = (0, 1, 0, 0), a
rotation by pi that is its own inverse, so they cannot tell q from q^-1.
These tests pin the oracle's quaternion algebra, world_to_cam, cam_to_screen
and get_gaussian_tile_overlap independently of it:

- the reference's known answers: i j = k (tests/test_quaternion.mojo:10-20),
  (1, 0, 0) rotated 90 degrees about z is (0, 1, 0) by rotate_vec and by the
  rotation matrix, det R = 1 (:22-40), and the pinhole projection of
  (0, 0, 10) and (2, 0, 10) to u = 640 and u = 800 (tests/test_camera.mojo:11-34);
- scipy's rotation library in float64 on random rotations far from pi:
  quaternion product = matrix product, rotate_vec = R v, world_to_cam =
  R^T (p - eye) for the camera-to-world q, and the rasterizer's projected
  centres, depths and inverse screen covariances (J R^T Sigma R J^T + 0.3 I);
- cam_to_screen against f x / z + c in float64, including the |z| < 1e-4 clamp;
- the tile rectangle against a floor-and-clamp formulation in float64 (the
  reference truncates toward zero with Int(); after the pipeline's screen-box
  cull the two agree), on hand-derived boundary cases and on every visible
  splat of the 100k-splat 1280x720 workload, with the keys' emission order
  rebuilt from the rectangles.

A q <-> q^-1 swap, a transposed view rotation, ceil for trunc or a missed
clamp in the oracle each fail one of them.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from lumen_amd.camera import CameraConfig, PinholeCamera, look_at
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import random_scene, workload


def f32(*v):
    return np.ascontiguousarray(np.array(v, np.float32).ravel())


def out(n):
    return np.zeros(n, np.float32)


def qmul(oracle, a, b):
    r = out(4)
    oracle.lib().orc_quat_mul(oracle.fptr(f32(*a)), oracle.fptr(f32(*b)), oracle.fptr(r))
    return r


def scipy_rot(q_wxyz):
    q = np.asarray(q_wxyz, np.float64)
    return Rotation.from_quat([q[1], q[2], q[3], q[0]])


def random_rotations(rng, n, min_angle=0.3):
    """Unit quaternions (w, x, y, z) whose rotation angle lies in
    [min_angle, pi - min_angle]: far from the identity and from pi, where
    q and q^-1 give the same rotation."""
    qs = []
    while len(qs) < n:
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        ang = 2 * math.acos(min(1.0, abs(q[0])))
        if min_angle <= ang <= math.pi - min_angle:
            qs.append(q)
    return np.array(qs)


# ---- reference known answers ------------------------------------------------

def test_hamilton_product_kat(oracle):
    # tests/test_quaternion.mojo:10-20: i * j = k
    np.testing.assert_array_equal(qmul(oracle, (0, 1, 0, 0), (0, 0, 1, 0)), f32(0, 0, 0, 1))
    # the rest of the unit table, exact
    np.testing.assert_array_equal(qmul(oracle, (0, 0, 1, 0), (0, 1, 0, 0)), f32(0, 0, 0, -1))
    np.testing.assert_array_equal(qmul(oracle, (0, 0, 1, 0), (0, 0, 0, 1)), f32(0, 1, 0, 0))
    np.testing.assert_array_equal(qmul(oracle, (0, 0, 0, 1), (0, 1, 0, 0)), f32(0, 0, 1, 0))
    np.testing.assert_array_equal(qmul(oracle, (0, 1, 0, 0), (0, 1, 0, 0)), f32(-1, 0, 0, 0))


def from_axis_angle_f32(oracle, axis, angle):
    """Quatf.from_axis_angle (src/math/quaternion.mojo:34-44) in float32."""
    a = np.asarray(axis, np.float32)
    a = a / np.float32(np.sqrt(np.sum(a * a, dtype=np.float32)))
    half = np.float32(angle) * np.float32(0.5)
    s = np.sin(half, dtype=np.float32)
    q = f32(np.cos(half, dtype=np.float32), a[0] * s, a[1] * s, a[2] * s)
    r = out(4)
    oracle.lib().orc_quat_normalized(oracle.fptr(q), oracle.fptr(r))
    return r


def test_vector_rotation_kat(oracle):
    # tests/test_quaternion.mojo:22-40: (1, 0, 0) by 90 degrees about z -> (0, 1, 0)
    q = from_axis_angle_f32(oracle, (0, 0, 1), np.float32(math.pi / 2))
    v = out(3)
    oracle.lib().orc_quat_rotate_vec(oracle.fptr(q), oracle.fptr(f32(1, 0, 0)), oracle.fptr(v))
    np.testing.assert_allclose(v, [0, 1, 0], atol=1e-5)
    m = out(9)
    oracle.lib().orc_quat_to_matrix(oracle.fptr(q), oracle.fptr(m))
    R = m.reshape(3, 3).astype(np.float64)
    np.testing.assert_allclose(R @ [1, 0, 0], [0, 1, 0], atol=1e-5)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-5)


def projection_camera(pos, rot, W=1280, H=720, f=800.0):
    return PinholeCamera(CameraConfig(W, H, f, f, W / 2.0, H / 2.0, 0.01, 1000.0),
                         np.asarray(pos, np.float32), np.asarray(rot, np.float32))


def world_to_screen(oracle, cam, p):
    c = oracle.camera_struct(cam)
    pc, uv = out(3), out(2)
    oracle.lib().orc_world_to_cam(c, oracle.fptr(f32(*p)), oracle.fptr(pc))
    oracle.lib().orc_cam_to_screen(c, oracle.fptr(pc), oracle.fptr(uv))
    return pc, uv


def test_camera_projection_kat(oracle):
    # tests/test_camera.mojo:11-34: camera at the origin, identity rotation
    cam = projection_camera((0, 0, 0), (1, 0, 0, 0))
    _, uv = world_to_screen(oracle, cam, (0, 0, 10))
    np.testing.assert_allclose(uv, [640, 360], atol=1e-4)
    _, uv = world_to_screen(oracle, cam, (2, 0, 10))
    np.testing.assert_allclose(uv, [800, 360], atol=1e-4)


# ---- closed forms on general rotations ---------------------------------------

def test_quaternion_algebra_matches_rotation_library(oracle):
    rng = np.random.default_rng(11)
    qa, qb = random_rotations(rng, 500), random_rotations(rng, 500)
    for a, b in zip(qa, qb):
        ab = qmul(oracle, a, b).astype(np.float64)
        # q_a q_b is the rotation R_a R_b (the product order is not symmetric)
        np.testing.assert_allclose(scipy_rot(ab).as_matrix(),
                                   scipy_rot(a).as_matrix() @ scipy_rot(b).as_matrix(), atol=2e-6)
        m = out(9)
        oracle.lib().orc_quat_to_matrix(oracle.fptr(f32(*a)), oracle.fptr(m))
        np.testing.assert_allclose(m.reshape(3, 3), scipy_rot(a).as_matrix(), atol=2e-6)
        v = rng.normal(0, 3, 3)
        r = out(3)
        oracle.lib().orc_quat_rotate_vec(oracle.fptr(f32(*a)), oracle.fptr(f32(*v)), oracle.fptr(r))
        np.testing.assert_allclose(r, scipy_rot(a).apply(v), atol=3e-6 * np.linalg.norm(v) + 1e-7)
        inv = out(4)
        s = rng.uniform(0.5, 2.0)  # unnormalized: inverse = conjugate / |q|^2
        oracle.lib().orc_quat_inverse(oracle.fptr(f32(*(s * a))), oracle.fptr(inv))
        np.testing.assert_allclose(inv, np.array([a[0], -a[1], -a[2], -a[3]]) / s, rtol=2e-6,
                                   atol=2e-7)


def test_world_to_cam_is_inverse_camera_rotation(oracle):
    """p_cam = R(q)^T (p - eye) for the camera-to-world rotation q
    (src/scene/camera.mojo:69-71), on 1,000 rotations far from pi, with the
    quaternion unnormalized (the camera normalizes it, camera.mojo:40)."""
    rng = np.random.default_rng(12)
    qs = random_rotations(rng, 1000)
    worst = 0.0
    for q in qs:
        eye = rng.normal(0, 4, 3)
        p = rng.normal(0, 4, 3)
        cam = projection_camera(eye, q * rng.uniform(0.3, 3.0))
        pc, _ = world_to_screen(oracle, cam, p)
        d = np.asarray(p, np.float32).astype(np.float64) - np.asarray(eye, np.float32)
        want = scipy_rot(q).inv().apply(d)
        err = np.max(np.abs(pc - want)) / np.linalg.norm(d)
        worst = max(worst, err)
        # R(q) instead of R(q)^T (q for q^-1) lands far away
        assert np.max(np.abs(scipy_rot(q).apply(d) - want)) / np.linalg.norm(d) > 1e-2
    assert worst < 2e-6, worst


def test_cam_to_screen_closed_form(oracle):
    rng = np.random.default_rng(13)
    cam = PinholeCamera(CameraConfig(1920, 1080, 1400.0, 1350.0, 955.5, 541.25, 0.01, 1000.0),
                        np.zeros(3, np.float32), np.array([1, 0, 0, 0], np.float32))
    c = oracle.camera_struct(cam)
    zs = np.concatenate([rng.uniform(0.01, 50, 500), -rng.uniform(0.01, 50, 100),
                         [5e-5, -5e-5, 0.0, 1e-4, -1e-4, 2e-4, -2e-4, 9.9e-5]])
    for z in zs.astype(np.float32):
        x, y = rng.normal(0, 3, 2).astype(np.float32)
        uv = out(2)
        oracle.lib().orc_cam_to_screen(c, oracle.fptr(f32(x, y, z)), oracle.fptr(uv))
        tz = 1e-4 if (-1e-4 < z < 1e-4) else float(z)
        want = np.array([1400.0 * x / tz + 955.5, 1350.0 * y / tz + 541.25])
        np.testing.assert_allclose(uv, want, rtol=2e-6, atol=2e-4)


def test_rasterizer_projection_off_axis(oracle):
    """The projected records of orc_rasterize against the float64 camera
    model with a general (non-pi) look-at rotation: centres, depth and the
    inverse EWA screen covariance J W Sigma W^T J^T + 0.3 I with W = R(q)^T."""
    sc = random_scene(20_000, 1.0, 21)
    cam = look_at(CameraConfig(640, 480, 600.0, 600.0, 320.0, 240.0, 0.01, 1000.0), (1.3, 3.9, 2.8))
    q = np.asarray(cam.rotation, np.float64)
    assert 0.3 < 2 * math.acos(min(1.0, abs(q[0]) / np.linalg.norm(q))) < math.pi - 0.3
    ref = oracle.rasterize(sc, cam, RenderConfig())
    V = ref["visible"]
    assert V > 5000
    idx = ref["vis_to_orig"]
    P = ref["proj"].astype(np.float64)
    pos = sc.planes[0:3, idx].T.astype(np.float64)
    Rc = scipy_rot(q).as_matrix()
    t = (pos - np.asarray(cam.position, np.float32).astype(np.float64)) @ Rc  # rows: R^T (p - eye)
    f, c = 600.0, (320.0, 240.0)
    u = f * t[:, 0] / t[:, 2] + c[0]
    v = f * t[:, 1] / t[:, 2] + c[1]
    np.testing.assert_allclose(P[:, 0], u, rtol=0, atol=2e-3)
    np.testing.assert_allclose(P[:, 1], v, rtol=0, atol=2e-3)
    np.testing.assert_allclose(P[:, 2], t[:, 2], rtol=2e-6)
    # screen covariance
    s = np.maximum(sc.planes[3:6, idx].T.astype(np.float64), 1e-5)
    qg = sc.planes[6:10, idx].T.astype(np.float64)
    Rg = Rotation.from_quat(qg[:, [1, 2, 3, 0]]).as_matrix()
    S3 = Rg @ (s[:, :, None] ** 2 * np.transpose(Rg, (0, 2, 1)))
    tz = t[:, 2]
    J = np.zeros((V, 2, 3))
    J[:, 0, 0] = f / tz
    J[:, 0, 2] = -f * t[:, 0] / tz ** 2
    J[:, 1, 1] = f / tz
    J[:, 1, 2] = -f * t[:, 1] / tz ** 2
    T = J @ Rc.T
    S2 = T @ S3 @ np.transpose(T, (0, 2, 1)) + 0.3 * np.eye(2)
    inv = np.linalg.inv(S2)
    got = np.stack([P[:, 3], P[:, 4], P[:, 4], P[:, 5]], axis=1).reshape(V, 2, 2)
    err = np.max(np.abs(got - inv), axis=(1, 2)) / np.max(np.abs(inv), axis=(1, 2))
    cond = np.linalg.cond(S2)
    ok = cond < 1e3
    assert ok.sum() > 0.9 * V
    assert np.max(err[ok]) < 1e-3, np.max(err[ok])
    # with W = R(q) (the transposition mistake) the covariances move
    T_bad = J @ Rc
    S2_bad = T_bad @ S3 @ np.transpose(T_bad, (0, 2, 1)) + 0.3 * np.eye(2)
    bad = np.max(np.abs(np.linalg.inv(S2_bad) - inv), axis=(1, 2)) / np.max(np.abs(inv), axis=(1, 2))
    assert np.median(bad) > 1e-2


# ---- tile rectangles ---------------------------------------------------------

def floor_rect(cx, cy, r, W, H, tw, th):
    """get_gaussian_tile_overlap (src/raster/tile.mojo:62-97) as a floor and
    clamp in float64 over the float32 box edges c -+ r: the pixel range is
    [floor(max(c - r, 0)), floor(min(c + r, size - 1))] and tiles are its ends
    divided by the tile size.  Returns (ok, min_tx, min_ty, max_tx, max_ty)
    arrays; ok is False for boxes that miss the image."""
    cx, cy = np.asarray(cx, np.float32), np.asarray(cy, np.float32)
    rf = np.asarray(r).astype(np.float32)
    x0, x1 = (cx - rf).astype(np.float64), (cx + rf).astype(np.float64)
    y0, y1 = (cy - rf).astype(np.float64), (cy + rf).astype(np.float64)
    ok = (x1 >= 0) & (x0 < W) & (y1 >= 0) & (y0 < H)
    return (ok,
            np.floor(np.clip(x0, 0, W - 1) / tw).astype(np.int64),
            np.floor(np.clip(y0, 0, H - 1) / th).astype(np.int64),
            np.floor(np.clip(x1, 0, W - 1) / tw).astype(np.int64),
            np.floor(np.clip(y1, 0, H - 1) / th).astype(np.int64))


def orc_rect(oracle, W, H, tw, th, cx, cy, r):
    o = np.zeros(4, np.int64)
    ok = oracle.lib().orc_tile_overlap(W, H, tw, th, cx, cy, r,
                                       o.ctypes.data_as(oracle.i64p))
    return ok, tuple(int(v) for v in o)


@pytest.mark.parametrize("cx,cy,r,want", [
    (18.5, 18.5, 3, (0, 0, 1, 1)),          # 15.5 -> 15 (tile 0); ceil would give tile 1
    (-2.5, 100.0, 3, (0, 6, 0, 6)),         # straddles x = 0: min clamps to pixel 0
    (1279.5, 706.0, 3, (79, 43, 79, 44)),   # straddles x = W: max clamps to pixel 1279
    (1282.9, 10.0, 3, (79, 0, 79, 0)),      # 1279.9 -> pixel 1279, the last column
    (640.0, 360.0, 10000, (0, 0, 79, 44)),  # covers every tile
    (16.0, 16.0, 0, (1, 1, 1, 1)),          # a point on a tile corner
])
def test_tile_overlap_boundary_kats(oracle, cx, cy, r, want):
    ok, rect = orc_rect(oracle, 1280, 720, 16, 16, cx, cy, r)
    assert ok and rect == want
    fok, *frect = floor_rect(cx, cy, r, 1280, 720, 16, 16)
    assert bool(fok) and tuple(int(v) for v in frect) == want


@pytest.mark.parametrize("cx,cy,r", [(1283.0, 10.0, 3), (10.0, 723.0, 3), (-5.0, 10.0, 3),
                                     (10.0, -3.5, 3)])
def test_tile_overlap_rejects(oracle, cx, cy, r):
    ok, _ = orc_rect(oracle, 1280, 720, 16, 16, cx, cy, r)
    if (cx, cy) == (10.0, -3.5):
        # cy + r = -0.5: Int() truncates toward zero to pixel 0, so the
        # reference's tile test alone keeps it; the screen-box cull before it
        # (culling.mojo:53-71) rejects it in the pipeline
        assert ok
        assert not oracle.lib().orc_screen_box_visible(1280, 720, cx, cy, r)
    else:
        assert not ok


@pytest.mark.parametrize("tw,th", [(16, 16), (8, 8), (32, 8)])
def test_tile_rects_and_emission_of_the_100k_workload(oracle, tw, th):
    """Every visible splat's rectangle on the 100k-splat 1280x720 workload
    (splats straddle x = 0 and x = W) equals the floor-and-clamp rule, and
    the unsorted keys are the rectangles enumerated per visible splat,
    ty-major then tx (pipeline.mojo:141-157), with depth key trunc(z 1000)."""
    sc, cam = workload("random100k_1280x720")
    W, H = cam.config.width, cam.config.height
    ref = oracle.rasterize(sc, cam, RenderConfig(tile_width=tw, tile_height=th))
    P, rad, rect = ref["proj"], ref["radius"], ref["rect"]
    ok, x0, y0, x1, y1 = floor_rect(P[:, 0], P[:, 1], rad, W, H, tw, th)
    assert ok.all()
    np.testing.assert_array_equal(rect, np.stack([x0, y0, x1, y1], axis=1))
    rf = rad.astype(np.float32)
    assert ((P[:, 0] - rf) < 0).sum() > 50 and ((P[:, 0] + rf) >= W).sum() > 50
    # emission order
    tiles_x = (W + tw - 1) // tw
    keys = []
    for v in range(ref["visible"]):
        tys = np.arange(y0[v], y1[v] + 1)
        txs = np.arange(x0[v], x1[v] + 1)
        tile = (tys[:, None] * tiles_x + txs[None, :]).ravel().astype(np.uint64)
        d = np.float32(P[v, 2]) * np.float32(1000.0)
        dk = np.uint64(min(max(int(np.trunc(d)), 0), 0x7FFFFFFF))
        keys.append((tile << np.uint64(32)) | dk)
    keys = np.concatenate(keys)
    np.testing.assert_array_equal(ref["keys_unsorted"], keys)
    counts = (x1 - x0 + 1) * (y1 - y0 + 1)
    np.testing.assert_array_equal(ref["vals_unsorted"], np.repeat(np.arange(ref["visible"]), counts))
    assert ref["instances"] == counts.sum()
