"""The reference's known-answer tests, restated against the CPU oracle.

Each test names the reference test it restates.  Together with the closed
forms and invariants below they pin the oracle before it is trusted as the
checker of the HIP path.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, PinholeCamera, look_at
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import GaussianScene, lattice_scene


def f32(*v):
    return np.ascontiguousarray(np.array(v, np.float32).ravel())


def out(n):
    return np.zeros(n, np.float32)


# ---- tests/test_projection.mojo:12-59 -----------------------------------
def test_jacobian_analytical(oracle):
    J = out(6)
    oracle.lib().orc_jacobian(oracle.fptr(f32(0, 0, 10)), 1000.0, 1000.0, oracle.fptr(J))
    assert J[0] == pytest.approx(100.0, abs=1e-4) and J[4] == pytest.approx(100.0, abs=1e-4)
    assert J[2] == pytest.approx(0.0, abs=1e-4) and J[5] == pytest.approx(0.0, abs=1e-4)
    oracle.lib().orc_jacobian(oracle.fptr(f32(2, 3, 10)), 1000.0, 1000.0, oracle.fptr(J))
    assert J[2] == pytest.approx(-20.0, abs=1e-4)
    assert J[5] == pytest.approx(-30.0, abs=1e-4)


def test_ewa_screen_covariance(oracle):
    c2 = out(3)
    eye3 = f32(1, 0, 0, 0, 1, 0, 0, 0, 1)
    oracle.lib().orc_ewa(oracle.fptr(f32(1, 0, 0, 1, 0, 1)), oracle.fptr(f32(0, 0, 10)),
                         oracle.fptr(eye3), 100.0, 100.0, 0.3, oracle.fptr(c2))
    assert c2[0] == pytest.approx(100.3, abs=1e-3) and c2[2] == pytest.approx(100.3, abs=1e-3)
    assert c2[1] == pytest.approx(0.0, abs=1e-4)
    inv = out(3)
    oracle.lib().orc_inverse2d(oracle.fptr(c2), oracle.fptr(inv))
    assert c2[0] * inv[0] + c2[1] * inv[1] == pytest.approx(1.0, abs=1e-4)
    assert c2[1] * inv[1] + c2[2] * inv[2] == pytest.approx(1.0, abs=1e-4)


def test_ewa_matches_float64_closed_form(oracle):
    """Sigma2D = J W Sigma W^T J^T + 0.3 I for a generic pose (float64 closed form)."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        s = np.exp(rng.uniform(-3, 0, 3))
        cov = out(6)
        oracle.lib().orc_cov3d(oracle.fptr(f32(*s)), oracle.fptr(f32(*q)), oracle.fptr(cov))
        from lumen_amd.camera import quaternion_to_rotation
        R = quaternion_to_rotation(q)
        S = R @ np.diag(s ** 2) @ R.T
        np.testing.assert_allclose(cov, [S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]],
                                   atol=1e-6, rtol=1e-5)
        t = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(2, 6)])
        qw = rng.standard_normal(4)
        qw /= np.linalg.norm(qw)
        W = quaternion_to_rotation(qw)
        c2 = out(3)
        oracle.lib().orc_ewa(oracle.fptr(cov), oracle.fptr(f32(*t)), oracle.fptr(f32(*W.ravel())),
                             700.0, 650.0, 0.3, oracle.fptr(c2))
        fx, fy = 700.0, 650.0
        J = np.array([[fx / t[2], 0, -fx * t[0] / t[2] ** 2], [0, fy / t[2], -fy * t[1] / t[2] ** 2]])
        T = J @ W
        ref = T @ S @ T.T + 0.3 * np.eye(2)
        np.testing.assert_allclose(c2, [ref[0, 0], ref[0, 1], ref[1, 1]], rtol=2e-4, atol=2e-3)


# ---- tests/test_covariance.mojo:11-37 -----------------------------------
def test_covariance_axis_aligned(oracle):
    c = out(6)
    oracle.lib().orc_cov3d(oracle.fptr(f32(2, 3, 4)), oracle.fptr(f32(1, 0, 0, 0)), oracle.fptr(c))
    np.testing.assert_allclose(c, [4, 0, 0, 9, 0, 16], atol=1e-5)
    M = np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]], np.float64)
    assert np.linalg.det(M) == pytest.approx(576.0, abs=1e-3)


def test_covariance_rotation_invariance(oracle):
    half = math.pi / 4  # rotation of pi/2 about +y
    q = f32(math.cos(half), 0, math.sin(half), 0)
    c = out(6)
    oracle.lib().orc_cov3d(oracle.fptr(f32(1.5, 2.5, 3.5)), oracle.fptr(q), oracle.fptr(c))
    M = np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]], np.float64)
    assert np.linalg.det(M) == pytest.approx((1.5 * 2.5 * 3.5) ** 2, abs=1e-2)
    assert np.all(np.linalg.eigvalsh(M) > 0)


def test_scale_clamp_regression(oracle):
    """tests/test_regressions.mojo:26-57: scale (1,0,0) -> det > 0 after low-pass."""
    c = out(6)
    oracle.lib().orc_cov3d(oracle.fptr(f32(1, 0, 0)), oracle.fptr(f32(1, 0, 0, 0)), oracle.fptr(c))
    c2 = out(3)
    oracle.lib().orc_ewa(oracle.fptr(c), oracle.fptr(f32(0, 0, 2)),
                         oracle.fptr(f32(1, 0, 0, 0, 1, 0, 0, 0, 1)), 100.0, 100.0, 0.3,
                         oracle.fptr(c2))
    assert c2[0] * c2[2] - c2[1] * c2[1] > 0
    inv = out(3)
    oracle.lib().orc_inverse2d(oracle.fptr(c2), oracle.fptr(inv))
    assert inv[0] > 0 and inv[2] > 0


# ---- tests/test_cholesky.mojo:10-27 -------------------------------------
@pytest.mark.parametrize("cov,lmax,lmin,r", [((16, 0, 9), 16, 9, 12), ((2, 1, 2), 3, 1, 6)])
def test_eigen_radius(oracle, cov, lmax, lmin, r):
    lam = out(2)
    rad = oracle.lib().orc_eigen(oracle.fptr(f32(*cov)), oracle.fptr(lam))
    assert lam[0] == pytest.approx(lmax, abs=1e-5) and lam[1] == pytest.approx(lmin, abs=1e-5)
    assert rad == r


def test_eigen_radius_closed_form(oracle):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        a, c = rng.uniform(0.3, 400, 2)
        b = rng.uniform(-1, 1) * math.sqrt(a * c) * 0.99
        v = f32(a, b, c)
        lam = out(2)
        rad = oracle.lib().orc_eigen(oracle.fptr(v), oracle.fptr(lam))
        A = np.array([[v[0], v[1]], [v[1], v[2]]], np.float64)
        e = np.linalg.eigvalsh(A)
        assert lam[0] == pytest.approx(e[1], rel=1e-5)
        exact = 3.0 * math.sqrt(e[1])
        assert math.ceil(exact - 1e-3) <= rad <= math.ceil(exact + 1e-3)


# ---- tests/test_spherical_harmonics.mojo:10-47 --------------------------
def test_sh_basis_degrees(oracle):
    Y = out(16)
    oracle.lib().orc_sh_basis(oracle.fptr(f32(0, 0, 1)), oracle.fptr(Y))
    assert Y[0] == pytest.approx(0.2820948, abs=1e-5)
    assert Y[1] == pytest.approx(0.0, abs=1e-5) and Y[3] == pytest.approx(0.0, abs=1e-5)
    assert Y[2] == pytest.approx(0.4886025, abs=1e-5)
    oracle.lib().orc_sh_basis(oracle.fptr(f32(0, 0, -1)), oracle.fptr(Y))
    assert Y[2] == pytest.approx(-0.4886025, abs=1e-5)


def test_sh_color_evaluation(oracle):
    coeffs = np.zeros(48, np.float32)
    inv_y0 = np.float32(1.0) / np.float32(0.28209479)
    coeffs[0:3] = [(0.8 - 0.5) * inv_y0, (0.4 - 0.5) * inv_y0, (0.2 - 0.5) * inv_y0]
    col = out(3)
    oracle.lib().orc_sh_color(oracle.fptr(coeffs), oracle.fptr(f32(0, 1, 0)), 0, oracle.fptr(col))
    np.testing.assert_allclose(col, [0.8, 0.4, 0.2], atol=1e-4)


def _real_sh_reference(d):
    """Textbook real SH up to l = 3 with the 3DGS sign convention, float64."""
    x, y, z = d
    C0 = 0.5 * math.sqrt(1 / math.pi)
    C1 = 0.5 * math.sqrt(3 / math.pi)
    C20, C21, C22 = 0.5 * math.sqrt(15 / math.pi), 0.25 * math.sqrt(5 / math.pi), 0.25 * math.sqrt(15 / math.pi)
    C30, C31, C32 = 0.25 * math.sqrt(35 / (2 * math.pi)), 0.5 * math.sqrt(105 / math.pi), 0.25 * math.sqrt(21 / (2 * math.pi))
    C33, C34 = 0.25 * math.sqrt(7 / math.pi), 0.25 * math.sqrt(105 / math.pi)
    return [C0, -C1 * y, C1 * z, -C1 * x, C20 * x * y, -C20 * y * z, C21 * (2 * z * z - x * x - y * y),
            -C20 * x * z, C22 * (x * x - y * y), -C30 * y * (3 * x * x - y * y), C31 * x * y * z,
            -C32 * y * (4 * z * z - x * x - y * y), C33 * z * (2 * z * z - 3 * x * x - 3 * y * y),
            -C32 * x * (4 * z * z - x * x - y * y), C34 * z * (x * x - y * y),
            -C30 * x * (x * x - 3 * y * y)]


def test_sh_basis_matches_textbook(oracle):
    """Closed forms of the real SH (docs/MATHEMATICAL_FORMULATION.md:88-112)
    and orthonormality of the degree-3 basis under quadrature."""
    rng = np.random.default_rng(11)
    Y = out(16)
    for _ in range(200):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        oracle.lib().orc_sh_basis(oracle.fptr(f32(*d)), oracle.fptr(Y))
        np.testing.assert_allclose(Y, _real_sh_reference(d), atol=2e-6)
    # orthonormality: integral of Y_i Y_j over the sphere = delta_ij
    n = 20000
    u = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    pts = np.stack([np.sqrt(1 - u * u) * np.cos(phi), np.sqrt(1 - u * u) * np.sin(phi), u], 1)
    B = np.array([_real_sh_reference(p) for p in pts])
    G = (B.T @ B) * (4 * math.pi / n)
    np.testing.assert_allclose(G, np.eye(16), atol=0.06)


# ---- tests/test_culling.mojo:10-42, test_regressions.mojo:59-85 ---------
def test_frustum_depth_culling(oracle):
    L = oracle.lib()
    for z, vis in [(5.0, 1), (0.1, 1), (100.0, 1), (-1.0, 0), (0.0, 0), (0.05, 0), (150.0, 0)]:
        assert L.orc_depth_visible(np.float32(0.1), np.float32(100.0), z) == vis
    for z, vis in [(-5.0, 0), (0.2, 0), (5.0, 1)]:
        assert L.orc_depth_visible(0.5, 100.0, z) == vis


def test_frustum_screen_culling(oracle):
    L = oracle.lib()
    cases = [(400, 300, 1), (-10, 300, 1), (810, 300, 1), (-50, 300, 0), (900, 300, 0),
             (400, -50, 0), (400, 700, 0)]
    for cx, cy, vis in cases:
        assert L.orc_screen_box_visible(800, 600, cx, cy, 20) == vis


def test_opacity_cull_threshold(oracle):
    L = oracle.lib()
    assert L.orc_opacity_visible(np.float32(0.00392156862)) == 1
    assert L.orc_opacity_visible(np.nextafter(np.float32(0.00392156862), np.float32(0))) == 0


# ---- tests/test_radix_sort.mojo:14-68 -----------------------------------
def test_radix_sort_composite_keys(oracle):
    L = oracle.lib()
    keys = np.array([L.orc_encode_key(1, 5.0), L.orc_encode_key(0, 10.0),
                     L.orc_encode_key(0, 2.0), L.orc_encode_key(1, 1.0)], np.uint64)
    k, v = oracle.radix_sort(keys, np.arange(4))
    assert v.tolist() == [2, 1, 3, 0]
    assert (k >> np.uint64(32)).tolist() == [0, 0, 1, 1]


def lcg_keys(n=1000, seed=123456789):
    out_ = []
    s = seed
    for _ in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out_.append(s)
    return np.array(out_, np.uint64)


def test_radix_sort_lcg_matches_stable_sort(oracle):
    keys = lcg_keys()
    k, v = oracle.radix_sort(keys, np.arange(keys.size))
    assert np.all(k[:-1] <= k[1:])
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, order)


def test_radix_sort_stability_with_ties(oracle):
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 50, 20000).astype(np.uint64) << np.uint64(32) | \
        rng.integers(0, 30, 20000).astype(np.uint64)
    k, v = oracle.radix_sort(keys, np.arange(keys.size))
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(v, order)


def test_depth_key_encoding(oracle):
    L = oracle.lib()
    assert L.orc_encode_key(3, 1.0) == (3 << 32) | 1000
    assert L.orc_encode_key(0, -2.0) == 0
    assert L.orc_encode_key(7, 1e30) == (7 << 32) | 0x7FFFFFFF
    assert L.orc_encode_key(1, 0.0015) == (1 << 32) | 1


# ---- tests/test_binning.mojo:10-51 --------------------------------------
def test_tile_range_construction(oracle):
    L = oracle.lib()
    keys = np.array([L.orc_encode_key(t, d) for t, d in
                     [(0, 1), (0, 2), (2, 3), (2, 4), (2, 5), (3, 6), (3, 7)]], np.uint64)
    r = oracle.build_ranges(keys, 4)
    assert r.tolist() == [[0, 2], [-1, -1], [2, 5], [5, 7]]


def test_ranges_partition(oracle):
    rng = np.random.default_rng(9)
    keys = np.sort(rng.integers(0, 100, 5000).astype(np.uint64) << np.uint64(32))
    r = oracle.build_ranges(keys, 120)
    nonempty = r[r[:, 0] >= 0]
    assert nonempty[0, 0] == 0 and nonempty[-1, 1] == keys.size
    assert np.all(nonempty[1:, 0] == nonempty[:-1, 1])
    for t in range(120):
        assert (r[t, 1] - r[t, 0] if r[t, 0] >= 0 else 0) == np.sum((keys >> np.uint64(32)) == t)


# ---- tests/test_compositor.mojo:10-50 -----------------------------------
def _blend(oracle, colors, alphas, depths):
    st = f32(0, 0, 0, 1, 0)
    idx = oracle.lib().orc_blend_sequence(oracle.fptr(f32(*colors)), oracle.fptr(f32(*alphas)),
                                          oracle.fptr(f32(*depths)), len(alphas), oracle.fptr(st))
    return st, idx


def test_single_splat_compositing(oracle):
    st, idx = _blend(oracle, [1, 0, 0], [0.6], [5.0])
    assert idx == -1
    np.testing.assert_allclose(st, [0.6, 0, 0, 0.4, 3.0], atol=1e-5)


def test_two_splat_order(oracle):
    st, _ = _blend(oracle, [0, 1, 0, 0, 0, 1], [0.5, 0.8], [2.0, 4.0])
    assert st[1] == pytest.approx(0.5, abs=1e-5) and st[2] == pytest.approx(0.4, abs=1e-5)
    assert st[3] == pytest.approx(0.1, abs=1e-5)


def test_early_ray_termination(oracle):
    st, idx = _blend(oracle, [1, 1, 1] * 10, [0.9] * 10, [1.0] * 10)
    assert st[3] <= 1e-4
    # float32: 1 - 0.9f = 0.100000024, its 4th power is 1.000001e-4 > 1e-4,
    # so the ray terminates on the fifth splat
    t = np.float32(1.0)
    for k in range(10):
        t = np.float32(t * (np.float32(1.0) - np.float32(0.9)))
        if t <= np.float32(1e-4):
            break
    assert idx == k == 4


def test_min_alpha_skip_has_no_effect(oracle):
    st, idx = _blend(oracle, [1, 1, 1, 1, 0, 0], [0.0039, 0.5], [1.0, 2.0])
    np.testing.assert_allclose(st, [0.5, 0, 0, 0.5, 1.0], atol=1e-7)


def test_pixel_alpha_closed_form(oracle):
    L = oracle.lib()
    inv = f32(0.25, 0.0, 0.25)
    # power = 0.5 * (dx^2 + dy^2) / 4 at (dx, dy) = (2, 0) -> 0.5
    a = L.orc_pixel_alpha(2.5, 0.5, 0.5, 0.5, oracle.fptr(inv), 0.8, 0)
    assert a == pytest.approx(0.8 * math.exp(-0.5), rel=1e-6)
    assert L.orc_pixel_alpha(0.5, 0.5, 0.5, 0.5, oracle.fptr(inv), 1.0, 0) == np.float32(0.99)
    # power > 4.5 -> 0: dx = 6.1 -> power = 4.65
    assert L.orc_pixel_alpha(6.6, 0.5, 0.5, 0.5, oracle.fptr(inv), 1.0, 0) == 0.0


# ---- exp used by the compositor ------------------------------------------
def test_exp_accuracy(oracle):
    """orc_expf within 1 ulp of exp (float64) on the compositor's domain
    [-4.5, 0] (every 256th float there, and a uniform grid; over every float
    the maximum is 0.994 ulp) and within 2 ulp over the normal range."""
    bits = np.arange(0x80000000, 0xC0900001, 256, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    y = oracle.expf_array(x)
    ref = np.exp(x.astype(np.float64))
    assert np.max(np.abs(y - ref) / np.spacing(ref.astype(np.float32))) <= 1.0
    x = np.linspace(-4.5, 0.0, 400001, dtype=np.float32)
    y = oracle.expf_array(x)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(y - ref) / ulp) <= 1.0
    x = np.random.default_rng(0).uniform(-87, 88, 200000).astype(np.float32)
    y = oracle.expf_array(x)
    ref = np.exp(x.astype(np.float64))
    assert np.max(np.abs(y - ref) / np.spacing(ref.astype(np.float32))) <= 2.0
    assert oracle.expf(0.0) == 1.0 and oracle.expf(-0.0) == 1.0


def test_exp_polynomial_is_the_fit():
    """The coefficients in both exp sources are the float32 rounding of
    oracle/exp_poly_fit.py's minimax fit, and the fit's relative error is
    far below float32 resolution (2^-24 = 6e-8)."""
    import os
    import re
    from oracle.exp_poly_fit import fit
    c, err = fit()
    assert err < 3e-9
    want = [float(x) for x in c.astype(np.float32)[1:]][::-1]  # c6 .. c2, Horner order
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for rel, fn in (("oracle/lumen_oracle.c", "orc_expf"),
                    ("lumen_amd/csrc/splat_math.hpp", "splat_exp_neg")):
        src = open(os.path.join(root, rel)).read()
        body = src[src.index(fn + "(float x)"):]
        body = body[:body.index("ldexpf")]
        got = [float(np.float32(float(t[:-1])))
               for t in re.findall(r"p = (?:fmaf\(p, r, )?([0-9.e+-]+f)", body)]
        assert got[:5] == want, (rel, got, want)
        assert float(np.float32(c[0])) == 1.0


def test_literal_constants_round_once():
    """Every FP32 literal of the path rounds to the same float whether taken
    directly from its decimal text or through float64 (so Mojo's literal
    materialisation cannot change it)."""
    lits = ["0.00392156862", "0.0001", "0.3", "0.99", "4.5", "0.000001", "0.00001", "0.0000001",
            "1000.0", "0.28209479177387814", "0.4886025119029199", "1.0925484305920792",
            "0.31539156525252005", "0.5462742152960396", "0.5900435899266435",
            "2.890611442640554", "0.4570457994644658", "0.3731763325901154", "1.445305721320277"]
    for s in lits:
        via64 = np.float32(float(s))
        exact = Fraction(s)
        lo, hi = np.nextafter(via64, np.float32(-np.inf)), np.nextafter(via64, np.float32(np.inf))
        d0 = abs(Fraction(float(via64)) - exact)
        assert d0 <= abs(Fraction(float(lo)) - exact) and d0 <= abs(Fraction(float(hi)) - exact), s


# ---- tests/test_rasterizer.mojo:15-71 -----------------------------------
def red_splat_scene():
    sh = np.zeros((1, 48), np.float32)
    sh[0, 0:3] = [1.77245, -1.77245, -1.77245]
    return GaussianScene.from_arrays([[0, 0, 5]], [[0.5, 0.5, 0.5]], [[1, 0, 0, 0]], [1.0], sh)


def red_splat_camera():
    cfg = CameraConfig(64, 64, 50.0, 50.0, 32.0, 32.0, 0.1, 100.0)
    return PinholeCamera(cfg, [0, 0, 0], [1, 0, 0, 0])


def test_single_splat_render(oracle):
    r = oracle.rasterize(red_splat_scene(), red_splat_camera(), RenderConfig())
    assert r["visible"] == 1 and r["instances"] > 0
    c = r["rgb"][32, 32]
    assert c[0] > 0.8 and c[1] < 0.2 and c[2] < 0.2
    np.testing.assert_allclose(r["rgb"][0, 0], [0, 0, 0], atol=1e-4)
    assert 4.5 < r["depth"][32, 32] < 5.5
    # closed form: Sigma2D = 25.3 I -> r = ceil(3 sqrt(25.3)) = 16, alpha at the
    # centre clamps to 0.99 and the colour is 0.99 * clamp(0.5 + C0 * 1.77245)
    assert r["radius"][0] == 16
    red = min(1.0, 0.5 + np.float32(0.28209479177387814) * np.float32(1.77245))
    assert c[0] == pytest.approx(0.99 * red, abs=1e-6)


def test_empty_and_extreme_rasterize(oracle):
    """tests/test_regressions.mojo:115-146."""
    cfg = CameraConfig(32, 32, 20.0, 20.0, 16.0, 16.0, 0.1, 50.0)
    cam = PinholeCamera(cfg, [0, 0, 0], [1, 0, 0, 0])
    empty = GaussianScene(np.zeros((11, 0), np.float32), np.zeros((0, 48), np.float32))
    assert oracle.rasterize(empty, cam, RenderConfig())["visible"] == 0
    far = GaussianScene.from_arrays([[1e5, 1e5, 5]], [[0.1] * 3], [[1, 0, 0, 0]], [1.0],
                                    np.zeros((1, 48), np.float32))
    r = oracle.rasterize(far, cam, RenderConfig())
    assert r["visible"] == 0 and r["instances"] == 0 and np.all(r["rgb"] == 0)


# ---- invariants of the whole path ----------------------------------------
def test_bruteforce_compositor_invariant(oracle):
    """Per pixel, blending ALL visible splats in global (depth key, visible
    index) order with the same skip/termination rules gives the tiled image
    exactly: splats outside a tile's list have alpha = 0 on its pixels.  This
    checks sort + ranges + tile loop against a loop that has none of them."""
    sc = lattice_scene(1001)
    cam = look_at(CameraConfig(160, 120, 140.0, 140.0, 80.0, 60.0, 0.01, 1000.0), (0.3, 0.2, 4.0))
    r = oracle.rasterize(sc, cam, RenderConfig())
    P = r["proj"]
    depth_key = np.array([oracle.lib().orc_encode_key(0, z) for z in P[:, 2]], np.uint64)
    order = np.lexsort((np.arange(len(P)), depth_key))
    H, W = 120, 160
    ys, xs = np.mgrid[0:H, 0:W]
    pxc = xs.astype(np.float32) + np.float32(0.5)
    pyc = ys.astype(np.float32) + np.float32(0.5)
    C = np.zeros((H, W, 3), np.float32)
    D = np.zeros((H, W), np.float32)
    T = np.ones((H, W), np.float32)
    done = np.zeros((H, W), bool)
    half, two = np.float32(0.5), np.float32(2.0)
    for v in order:
        cx, cy, z, i00, i01, i11, cr, cg, cb, op = P[v]
        dx = pxc - cx
        dy = pyc - cy
        power = half * (i00 * dx * dx + (two * i01) * dx * dy + i11 * dy * dy)
        power = np.where(power < 0, np.float32(0), power).astype(np.float32)
        g = oracle.expf_array(-power)
        alpha = (op * g).astype(np.float32)
        alpha = np.where(power > 4.5, np.float32(0), alpha)
        alpha = np.minimum(alpha, np.float32(0.99))
        act = (~done) & (alpha >= np.float32(0.00392156862))
        w = (alpha * T).astype(np.float32)
        for k, col in enumerate((cr, cg, cb)):
            C[..., k] = np.where(act, C[..., k] + col * w, C[..., k])
        D = np.where(act, D + z * w, D)
        T = np.where(act, T * (np.float32(1) - alpha), T)
        done |= act & (T <= np.float32(1e-4))
    np.testing.assert_array_equal(C, r["rgb"])
    np.testing.assert_array_equal(D, r["depth"])


def test_image_independent_of_tile_shape(oracle):
    """Keys and ranges depend on the tile shape; the image does not."""
    sc, cam = lattice_scene(1001), look_at(CameraConfig(200, 150, 175.0, 175.0, 100.0, 75.0), (0, 0, 4))
    base = oracle.rasterize(sc, cam, RenderConfig())
    for tw, th in [(8, 8), (32, 8)]:
        r = oracle.rasterize(sc, cam, RenderConfig(tile_width=tw, tile_height=th))
        np.testing.assert_array_equal(r["rgb"], base["rgb"])
        np.testing.assert_array_equal(r["depth"], base["depth"])


def test_exp_choice_within_tolerance(oracle):
    """libm expf instead of the path's exp moves no pixel by more than 1e-4
    (the tolerance that absorbs Mojo's own exp)."""
    sc, cam = lattice_scene(1001), look_at(CameraConfig(400, 300, 350.0, 350.0, 200.0, 150.0), (0, 0, 4))
    a = oracle.rasterize(sc, cam, RenderConfig(), exp_mode=0)
    b = oracle.rasterize(sc, cam, RenderConfig(), exp_mode=1)
    assert np.max(np.abs(a["rgb"] - b["rgb"])) <= 1e-4
    assert np.array_equal(a["keys"], b["keys"])


# ---- tests/test_loss.mojo:9-73 ------------------------------------------
def test_l1_loss_and_gradients(oracle):
    L = oracle.lib()
    pred, target, grad = f32(0, 0, 0, 0), f32(0, 0, 0, 0), out(4)
    assert L.orc_l1_loss(oracle.fptr(pred), oracle.fptr(target), 4, oracle.fptr(grad)) == 0.0
    pred, target = f32(0.8, 0.2, 0.5, 0.1), f32(0.5, 0.5, 0.5, 0.5)
    v = L.orc_l1_loss(oracle.fptr(pred), oracle.fptr(target), 4, oracle.fptr(grad))
    assert v == pytest.approx(0.25, abs=1e-5)
    np.testing.assert_allclose(grad, [0.25, -0.25, 0.25, -0.25], atol=1e-5)


def test_psnr_scaling(oracle):
    L = oracle.lib()
    t = np.zeros(100, np.float32)
    assert L.orc_psnr(oracle.fptr(t), oracle.fptr(t), 100) > 90
    assert L.orc_psnr(oracle.fptr(np.full(100, 0.1, np.float32)), oracle.fptr(t), 100) == pytest.approx(20, abs=0.1)
    assert L.orc_psnr(oracle.fptr(np.full(100, 0.01, np.float32)), oracle.fptr(t), 100) == pytest.approx(40, abs=0.1)


def test_ssim_identical(oracle):
    a = np.full(16 * 16 * 3, 0.5, np.float32)
    assert oracle.lib().orc_ssim(oracle.fptr(a), oracle.fptr(a.copy()), 16, 16, 8) == pytest.approx(1.0, abs=1e-4)


def test_power_limit_is_conservative(oracle):
    """The compositors skip a pixel before the exp when power > min(4.5,
    ln(255 op) + 1e-3): there the reference's alpha = op * exp(-power) is below
    1/255 (or the power test applies), so the skip changes nothing.  Checked
    with the oracle's exp at the first float above the limit and beyond."""
    rng = np.random.default_rng(77)
    op = np.concatenate([rng.uniform(1.0 / 255.0, 1.0, 200_000),
                         np.float32(1.0 / 255.0) * (1 + rng.uniform(0, 1e-3, 20_000)),
                         rng.uniform(0.34, 0.36, 20_000)]).astype(np.float32)
    lim = (np.log(np.float32(255.0) * op).astype(np.float32) + np.float32(0.001)).astype(np.float32)
    checked = 0
    for step in (0.0, 1e-6, 1e-4, 1e-2, 0.5):
        p = np.nextafter((lim + np.float32(step)).astype(np.float32), np.float32(np.inf))
        keep = p <= np.float32(4.5)  # beyond 4.5 the power test itself skips
        alpha = (op[keep] * oracle.expf_array(-p[keep].astype(np.float32))).astype(np.float32)
        assert np.all(alpha < np.float32(0.00392156862)), step
        checked += int(keep.sum())
    assert checked > 100_000


def test_limit_box_is_conservative():
    """The compositors skip a splat for a wave whose pixel centres all lie
    outside its box; with the power limit the box half-widths are
    (h - 0.02) sqrt(lim / 4.5) + 0.02 of the 3-sigma box h = 3 sqrt(i11 / det)
    1.02 + 0.02.  Just outside that box the power, computed in float32 as the
    compositor does, exceeds the limit even at the row that minimises it."""
    rng = np.random.default_rng(78)
    f32 = np.float32
    n = 200_000
    # inverse covariances with condition numbers up to ~1e4 (the box's validity range)
    lam1 = np.exp(rng.uniform(np.log(1e-4), np.log(10.0), n))
    lam2 = lam1 * np.exp(rng.uniform(0.0, np.log(9e3), n))
    th = rng.uniform(0, np.pi, n)
    c, s = np.cos(th), np.sin(th)
    i00 = (lam1 * c * c + lam2 * s * s).astype(f32)
    i11 = (lam1 * s * s + lam2 * c * c).astype(f32)
    i01 = ((lam1 - lam2) * c * s).astype(f32)
    det = (i00 * i11 - i01 * i01).astype(f32)
    tr = (i00 + i11).astype(f32)
    ok = (i00 > 0) & (i11 > 0) & (det > 0) & (tr * tr <= f32(1e4) * det)
    i00, i11, i01, det = i00[ok], i11[ok], i01[ok], det[ok]
    op = rng.uniform(1.0 / 255.0, 1.0, i00.size).astype(f32)
    lim = np.minimum((np.log(f32(255.0) * op).astype(f32) + f32(0.001)).astype(f32), f32(4.5))
    hx = (f32(3.0) * np.sqrt(i11 / det).astype(f32) * f32(1.02) + f32(0.02)).astype(f32)
    s_ = np.sqrt(lim * f32(1.0 / 4.5)).astype(f32)
    hxl = ((hx - f32(0.02)) * s_ + f32(0.02)).astype(f32)
    dx = np.nextafter(hxl, np.float32(np.inf)).astype(f32)
    a2 = (f32(2.0) * i01).astype(f32)
    for k in (-1, 0, 1):
        dy = (-(i01.astype(np.float64) * dx) / i11 + k * 1e-3).astype(f32)
        power = (f32(0.5) * ((i00 * dx * dx + a2 * dx * dy).astype(f32) + (i11 * dy * dy))).astype(f32)
        assert np.all(power > lim), k
