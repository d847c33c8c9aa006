"""ctypes binding of the CPU oracle (oracle/liblumen_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / CPU baseline.  The product path
never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liblumen_oracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "liblumen_ref.so")

f32p = ctypes.POINTER(ctypes.c_float)
i64p = ctypes.POINTER(ctypes.c_int64)
u64p = ctypes.POINTER(ctypes.c_uint64)
i32p = ctypes.POINTER(ctypes.c_int32)


class OrcCamera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int64), ("height", ctypes.c_int64),
                ("focal_x", ctypes.c_float), ("focal_y", ctypes.c_float),
                ("principal_x", ctypes.c_float), ("principal_y", ctypes.c_float),
                ("near_plane", ctypes.c_float), ("far_plane", ctypes.c_float),
                ("pos", ctypes.c_float * 3), ("rot", ctypes.c_float * 4)]


class OrcConfig(ctypes.Structure):
    _fields_ = [("tile_width", ctypes.c_int64), ("tile_height", ctypes.c_int64),
                ("lowpass_filter", ctypes.c_float), ("max_sh_degree", ctypes.c_int64),
                ("white_background", ctypes.c_int32), ("compute_depth", ctypes.c_int32),
                ("exp_mode", ctypes.c_int32)]


class OrcScene(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [(k, f32p) for k in (
        "pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z", "rot_w", "rot_x", "rot_y",
        "rot_z", "opacity", "sh")]


class OrcResult(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in (
        "n", "visible", "instances", "tiles", "tiles_x", "tiles_y", "evaluations", "width",
        "height")] + [
        ("vis_to_orig", i64p), ("proj", f32p), ("radius", i64p), ("rect", i64p),
        ("keys_unsorted", u64p), ("vals_unsorted", i64p), ("keys", u64p), ("vals", i64p),
        ("ranges", i64p), ("rgb", f32p), ("depth", f32p), ("t_final", f32p), ("n_eval", i32p)]


class OrcLegacyCamera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("target", ctypes.c_float * 3),
                ("up", ctypes.c_float * 3), ("fov_x_deg", ctypes.c_float),
                ("fov_y_deg", ctypes.c_float), ("near_plane", ctypes.c_float),
                ("far_plane", ctypes.c_float)]


class OrcLegacyConfig(ctypes.Structure):
    _fields_ = [("near_plane", ctypes.c_float), ("far_plane", ctypes.c_float),
                ("early_termination_alpha", ctypes.c_float),
                ("gaussian_radius_sigma", ctypes.c_float), ("tile_size", ctypes.c_uint32)]


class OrcLegacyStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("total", "visible", "culled", "total_tiles",
                                               "touched")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_PATH)
        L.orc_rasterize.restype = ctypes.POINTER(OrcResult)
        L.orc_rasterize.argtypes = [ctypes.POINTER(OrcScene), ctypes.POINTER(OrcCamera),
                                    ctypes.POINTER(OrcConfig)]
        L.orc_result_free.argtypes = [ctypes.POINTER(OrcResult)]
        L.orc_expf.restype = ctypes.c_float
        L.orc_expf.argtypes = [ctypes.c_float]
        L.orc_expf_array.restype = None
        L.orc_expf_array.argtypes = [f32p, f32p, ctypes.c_int64]
        for name, args in {
            "orc_quat_normalized": [f32p, f32p], "orc_quat_to_matrix": [f32p, f32p],
            "orc_quat_inverse": [f32p, f32p], "orc_quat_mul": [f32p, f32p, f32p],
            "orc_quat_rotate_vec": [f32p, f32p, f32p],
            "orc_world_to_cam": [ctypes.POINTER(OrcCamera), f32p, f32p],
            "orc_cam_to_screen": [ctypes.POINTER(OrcCamera), f32p, f32p],
            "orc_cov3d": [f32p, f32p, f32p],
            "orc_jacobian": [f32p, ctypes.c_float, ctypes.c_float, f32p],
            "orc_ewa": [f32p, f32p, f32p, ctypes.c_float, ctypes.c_float, ctypes.c_float, f32p],
            "orc_inverse2d": [f32p, f32p], "orc_sh_basis": [f32p, f32p],
            "orc_sh_color": [f32p, f32p, ctypes.c_int64, f32p],
            "orc_radix_sort": [u64p, i64p, ctypes.c_int64],
            "orc_build_ranges": [u64p, ctypes.c_int64, ctypes.c_int64, i64p],
        }.items():
            getattr(L, name).restype = None
            getattr(L, name).argtypes = args
        L.orc_eigen.restype = ctypes.c_int64
        L.orc_eigen.argtypes = [f32p, f32p]
        L.orc_depth_visible.restype = ctypes.c_int
        L.orc_depth_visible.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_opacity_visible.restype = ctypes.c_int
        L.orc_opacity_visible.argtypes = [ctypes.c_float]
        L.orc_screen_box_visible.restype = ctypes.c_int
        L.orc_screen_box_visible.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_int64]
        L.orc_tile_overlap.restype = ctypes.c_int
        L.orc_tile_overlap.argtypes = [ctypes.c_int64] * 4 + [ctypes.c_float, ctypes.c_float,
                                                              ctypes.c_int64, i64p]
        L.orc_encode_key.restype = ctypes.c_uint64
        L.orc_encode_key.argtypes = [ctypes.c_uint32, ctypes.c_float]
        L.orc_pixel_alpha.restype = ctypes.c_float
        L.orc_pixel_alpha.argtypes = [ctypes.c_float] * 4 + [f32p, ctypes.c_float, ctypes.c_int32]
        L.orc_blend_sequence.restype = ctypes.c_int64
        L.orc_blend_sequence.argtypes = [f32p, f32p, f32p, ctypes.c_int64, f32p]
        L.orc_l1_loss.restype = ctypes.c_float
        L.orc_l1_loss.argtypes = [f32p, f32p, ctypes.c_int64, f32p]
        L.orc_ssim.restype = ctypes.c_float
        L.orc_ssim.argtypes = [f32p, f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
        L.orc_psnr.restype = ctypes.c_float
        L.orc_psnr.argtypes = [f32p, f32p, ctypes.c_int64]
        L.orc_legacy_rasterize.restype = ctypes.c_int
        L.orc_legacy_rasterize.argtypes = [f32p, ctypes.c_int64, ctypes.POINTER(OrcLegacyCamera),
                                           ctypes.POINTER(OrcLegacyConfig), ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8), f32p,
                                           f32p, ctypes.POINTER(OrcLegacyStats)]
        L.orc_combined_loss.restype = ctypes.c_float
        L.orc_combined_loss.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        _lib = L
    return _lib


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(f32p)


def f32(*vals) -> np.ndarray:
    return np.ascontiguousarray(np.array(vals, np.float32).ravel())


def camera_struct(camera) -> OrcCamera:
    c = OrcCamera()
    cfg = camera.config
    c.width, c.height = cfg.width, cfg.height
    c.focal_x, c.focal_y = cfg.focal_x, cfg.focal_y
    c.principal_x, c.principal_y = cfg.principal_x, cfg.principal_y
    c.near_plane, c.far_plane = cfg.near_plane, cfg.far_plane
    for k in range(3):
        c.pos[k] = float(camera.position[k])
    for k in range(4):
        c.rot[k] = float(camera.rotation[k])
    return c


def config_struct(config, exp_mode: int = 0) -> OrcConfig:
    c = OrcConfig()
    c.tile_width, c.tile_height = config.tile_width, config.tile_height
    c.lowpass_filter = config.lowpass_filter
    c.max_sh_degree = config.max_sh_degree
    c.white_background = int(bool(config.white_background))
    c.compute_depth = int(bool(config.compute_depth))
    c.exp_mode = exp_mode
    return c


def rasterize(scene, camera, config, exp_mode: int = 0) -> dict:
    """Run the oracle's rasterize_scene; returns numpy copies of every stage."""
    L = lib()
    planes = np.ascontiguousarray(scene.planes, np.float32)
    sh = np.ascontiguousarray(scene.sh, np.float32)
    n = planes.shape[1]
    sc = OrcScene()
    sc.n = n
    names = ("pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z", "rot_w", "rot_x",
             "rot_y", "rot_z", "opacity")
    rows = [np.ascontiguousarray(planes[k]) for k in range(11)]
    for k, name in enumerate(names):
        setattr(sc, name, fptr(rows[k]))
    sc.sh = fptr(sh)
    cam = camera_struct(camera)
    cfg = config_struct(config, exp_mode)
    rp = L.orc_rasterize(ctypes.byref(sc), ctypes.byref(cam), ctypes.byref(cfg))
    r = rp.contents
    V, M, T = r.visible, r.instances, r.tiles
    W, H = r.width, r.height

    def arr(ptr, count, dtype):
        if count == 0:
            return np.zeros((0,), dtype)
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

    out = {
        "visible": V, "instances": M, "tiles": T, "tiles_x": r.tiles_x, "tiles_y": r.tiles_y,
        "evaluations": r.evaluations,
        "vis_to_orig": arr(r.vis_to_orig, V, np.int64),
        "proj": arr(r.proj, V * 10, np.float32).reshape(V, 10),
        "radius": arr(r.radius, V, np.int64),
        "rect": arr(r.rect, V * 4, np.int64).reshape(V, 4),
        "keys_unsorted": arr(r.keys_unsorted, M, np.uint64),
        "vals_unsorted": arr(r.vals_unsorted, M, np.int64),
        "keys": arr(r.keys, M, np.uint64),
        "vals": arr(r.vals, M, np.int64),
        "ranges": arr(r.ranges, 2 * T, np.int64).reshape(T, 2),
        "rgb": arr(r.rgb, W * H * 3, np.float32).reshape(H, W, 3),
        "depth": arr(r.depth, W * H, np.float32).reshape(H, W),
        "t_final": arr(r.t_final, W * H, np.float32).reshape(H, W),
        "n_eval": arr(r.n_eval, W * H, np.int32).reshape(H, W),
    }
    L.orc_result_free(rp)
    return out


def radix_sort(keys: np.ndarray, vals: np.ndarray):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.int64).copy()
    lib().orc_radix_sort(k.ctypes.data_as(u64p), v.ctypes.data_as(i64p), k.size)
    return k, v


def build_ranges(keys: np.ndarray, tiles: int) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint64)
    r = np.empty((tiles, 2), np.int64)
    lib().orc_build_ranges(k.ctypes.data_as(u64p), k.size, tiles, r.ctypes.data_as(i64p))
    return r


def expf(x: float) -> float:
    return lib().orc_expf(x)


def expf_array(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_expf_array(fptr(x), fptr(y), x.size)
    return y


# ---- FP64 forward / backward ---------------------------------------------
f64p = ctypes.POINTER(ctypes.c_double)


class OrcScene64(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("planes", f64p), ("sh", f64p)]


class OrcGrads64(ctypes.Structure):
    _fields_ = [("planes", f64p), ("sh", f64p), ("viewspace", f64p), ("count", i64p)]


def _lib64():
    L = lib()
    if not getattr(L, "_has64", False):
        L.orc64_forward.restype = ctypes.c_int
        L.orc64_forward.argtypes = [ctypes.POINTER(OrcScene64), ctypes.POINTER(OrcCamera),
                                    ctypes.POINTER(OrcConfig), f64p, f64p,
                                    ctypes.POINTER(ctypes.c_uint64)]
        L.orc64_backward.restype = ctypes.c_int
        L.orc64_backward.argtypes = [ctypes.POINTER(OrcScene64), ctypes.POINTER(OrcCamera),
                                     ctypes.POINTER(OrcConfig), f64p, f64p,
                                     ctypes.POINTER(OrcGrads64), f64p]
        L._has64 = True
    return L


def _scene64(planes, sh):
    planes = np.ascontiguousarray(planes, np.float64)
    sh = np.ascontiguousarray(sh, np.float64)
    s = OrcScene64()
    s.n = planes.shape[1]
    s.planes = planes.ctypes.data_as(f64p)
    s.sh = sh.ctypes.data_as(f64p)
    return s, (planes, sh)


def forward64(planes, sh, camera, config):
    """FP64 render: (rgb (H,W,3), depth (H,W), decision signature)."""
    L = _lib64()
    s, keep = _scene64(planes, sh)
    W, H = camera.config.width, camera.config.height
    rgb = np.zeros((H, W, 3), np.float64)
    depth = np.zeros((H, W), np.float64)
    sig = ctypes.c_uint64(0)
    L.orc64_forward(ctypes.byref(s), ctypes.byref(camera_struct(camera)),
                    ctypes.byref(config_struct(config)), rgb.ctypes.data_as(f64p),
                    depth.ctypes.data_as(f64p), ctypes.byref(sig))
    return rgb, depth, int(sig.value)


def backward64(planes, sh, camera, config, g_rgb, g_depth):
    """Analytic gradient of sum(g_rgb * rgb) + sum(g_depth * depth)."""
    L = _lib64()
    s, keep = _scene64(planes, sh)
    n = s.n
    gp = np.zeros((11, n), np.float64)
    gsh = np.zeros((n, 48), np.float64)
    vs = np.zeros(n, np.float64)
    cnt = np.zeros(n, np.int64)
    gr = OrcGrads64(gp.ctypes.data_as(f64p), gsh.ctypes.data_as(f64p), vs.ctypes.data_as(f64p),
                    cnt.ctypes.data_as(i64p))
    g_rgb = np.ascontiguousarray(g_rgb, np.float64)
    g_depth = np.ascontiguousarray(g_depth, np.float64)
    gvis = np.zeros((max(n, 1), 10), np.float64)
    V = L.orc64_backward(ctypes.byref(s), ctypes.byref(camera_struct(camera)),
                         ctypes.byref(config_struct(config)), g_rgb.ctypes.data_as(f64p),
                         g_depth.ctypes.data_as(f64p), ctypes.byref(gr),
                         gvis.ctypes.data_as(f64p))
    return {"planes": gp, "sh": gsh, "viewspace": vs, "count": cnt, "per_visible": gvis[:V]}


# ---- optimizer and density control (lumen_oracle_optim.c) --------------------

class OrcTrainConfig(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64)] + [(k, ctypes.c_float) for k in (
        "position_lr_init", "position_lr_final", "feature_lr", "opacity_lr", "scaling_lr",
        "rotation_lr", "beta1", "beta2", "epsilon", "densify_grad_threshold",
        "prune_opacity_threshold")]


class OrcAdamScalars(ctypes.Structure):
    _fields_ = [("pos_lr", ctypes.c_float), ("inv_b1", ctypes.c_float),
                ("inv_b2", ctypes.c_float)]


def train_config(cfg=None, **kw) -> OrcTrainConfig:
    """OrcTrainConfig from an object with TrainingConfig's fields (or the
    reference defaults) and keyword overrides."""
    d = dict(iterations=30000, position_lr_init=0.00016, position_lr_final=0.0000016,
             feature_lr=0.0025, opacity_lr=0.05, scaling_lr=0.005, rotation_lr=0.001, beta1=0.9,
             beta2=0.999, epsilon=1e-8, densify_grad_threshold=0.0002,
             prune_opacity_threshold=0.005)
    if cfg is not None:
        for k in d:
            if hasattr(cfg, k):
                d[k] = getattr(cfg, k)
    d.update(kw)
    c = OrcTrainConfig()
    for k, v in d.items():
        setattr(c, k, v)
    return c


def _optim_lib():
    L = lib()
    if not getattr(L, "_optim_ready", False):
        L.orc_adam_schedule.restype = None
        L.orc_adam_schedule.argtypes = [ctypes.POINTER(OrcTrainConfig), ctypes.c_int64,
                                        ctypes.POINTER(OrcAdamScalars)]
        L.orc_adam_step.restype = None
        L.orc_adam_step.argtypes = [f32p] * 8 + [ctypes.c_int64, ctypes.POINTER(OrcTrainConfig),
                                                 ctypes.POINTER(OrcAdamScalars), ctypes.c_int]
        L.orc_reset_opacities.restype = None
        L.orc_reset_opacities.argtypes = [f32p, ctypes.c_int64, ctypes.c_float]
        L.orc_densify_and_prune.restype = ctypes.c_int64
        L.orc_densify_and_prune.argtypes = [f32p, f32p, f32p, i32p, ctypes.c_int64,
                                            ctypes.POINTER(OrcTrainConfig), ctypes.c_float, f32p,
                                            f32p, ctypes.c_int64, i64p, i64p]
        L._optim_ready = True
    return L


def adam_schedule(cfg: OrcTrainConfig, step: int) -> OrcAdamScalars:
    s = OrcAdamScalars()
    _optim_lib().orc_adam_schedule(ctypes.byref(cfg), step, ctypes.byref(s))
    return s


def adam_step(planes, sh, gplanes, gsh, m, v, cfg: OrcTrainConfig, step: int,
              all_params: bool = False) -> None:
    """In place on float32 numpy arrays: planes (11, n), sh (n, 48), and the
    moments m = (m_planes, m_sh), v = (v_planes, v_sh)."""
    arrs = [planes, sh, gplanes, gsh, m[0], m[1], v[0], v[1]]
    for a in arrs:
        assert a.dtype == np.float32 and a.flags.c_contiguous
    s = adam_schedule(cfg, step)
    _optim_lib().orc_adam_step(*[fptr(a) for a in arrs], planes.shape[1], ctypes.byref(cfg),
                               ctypes.byref(s), int(bool(all_params)))


def reset_opacities(opacity: np.ndarray, max_val: float) -> None:
    assert opacity.dtype == np.float32 and opacity.flags.c_contiguous
    _optim_lib().orc_reset_opacities(fptr(opacity), opacity.size, max_val)


def densify_and_prune(planes, sh, vs_grad, vs_count, cfg: OrcTrainConfig, scene_extent: float):
    """(planes (11, m), sh (m, 48), origin (m,), (cloned, split, pruned))."""
    n = planes.shape[1]
    cap = max(2 * n, 1)
    op = np.zeros((11, cap), np.float32)
    osh = np.zeros((cap, 48), np.float32)
    origin = np.zeros(cap, np.int64)
    stats = np.zeros(3, np.int64)
    planes = np.ascontiguousarray(planes, np.float32)
    sh = np.ascontiguousarray(sh, np.float32)
    vs_grad = np.ascontiguousarray(vs_grad, np.float32)
    vs_count = np.ascontiguousarray(vs_count, np.int32)
    m = _optim_lib().orc_densify_and_prune(
        fptr(planes), fptr(sh), fptr(vs_grad), vs_count.ctypes.data_as(i32p), n,
        ctypes.byref(cfg), scene_extent, fptr(op), fptr(osh), cap,
        origin.ctypes.data_as(i64p), stats.ctypes.data_as(i64p))
    assert m >= 0
    return (np.ascontiguousarray(op[:, :m]), osh[:m].copy(), origin[:m].copy(),
            tuple(int(x) for x in stats))


def legacy_rasterize(aos, camera, config, rgb, depth, alpha):
    """The reference C library's lumen_rasterize (c_api/src/lumen_c.c:380-575)
    restated: aos (n, 59) float32 LumenGaussians; camera = dict(pos, target,
    up, fov_x, fov_y, near, far); config = dict(near, far, eta, sigma, tile);
    rgb (H, W, 3) uint8, depth and alpha (H, W) float32 framebuffer planes,
    updated in place.  Returns the stats as a dict."""
    aos = np.ascontiguousarray(aos, np.float32)
    H, W = depth.shape
    assert rgb.shape == (H, W, 3) and rgb.dtype == np.uint8 and rgb.flags["C_CONTIGUOUS"]
    assert alpha.shape == (H, W) and depth.flags["C_CONTIGUOUS"] and alpha.flags["C_CONTIGUOUS"]
    cam = OrcLegacyCamera()
    for k in range(3):
        cam.pos[k], cam.target[k], cam.up[k] = camera["pos"][k], camera["target"][k], camera["up"][k]
    cam.fov_x_deg, cam.fov_y_deg = camera["fov_x"], camera["fov_y"]
    cam.near_plane, cam.far_plane = camera["near"], camera["far"]
    cfg = OrcLegacyConfig(config["near"], config["far"], config["eta"], config["sigma"],
                          config["tile"])
    st = OrcLegacyStats()
    rc = lib().orc_legacy_rasterize(aos.ctypes.data_as(f32p) if aos.size else None,
                                    aos.shape[0] if aos.size else 0, ctypes.byref(cam),
                                    ctypes.byref(cfg), W, H,
                                    rgb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                    fptr(depth), fptr(alpha), ctypes.byref(st))
    if rc != 0:
        raise ValueError("orc_legacy_rasterize: invalid argument")
    return {k: int(getattr(st, k)) for k in ("total", "visible", "culled", "total_tiles",
                                             "touched")}
