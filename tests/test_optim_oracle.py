"""The optimizer oracle (oracle/lumen_oracle_optim.c) pinned on the CPU.

- densify_and_prune / reset_opacities: the reference's own known answers
  (tests/test_density_control.mojo:15-80), then the closed forms of the
  clone offset and the split children (density_control.mojo:76-91), the
  output order and the invariants the counts obey.
- Adam: a textbook float64 Adam (bias-corrected moments, Kingma & Ba
  Algorithm 1) over varying gradients must agree to float32 rounding; the
  first step moves each parameter by lr * g / (|g| + eps); the position rate
  follows the exponential schedule's end points (optimizer.mojo:103-110).
"""
import numpy as np
import pytest

from lumen_amd.scene import random_scene


def _scene(pos, scale, op, rot=None):
    n = len(pos)
    rot = rot if rot is not None else [(1.0, 0.0, 0.0, 0.0)] * n
    planes = np.concatenate([np.array(pos, np.float32).T, np.array(scale, np.float32).T,
                             np.array(rot, np.float32).T, np.array([op], np.float32)])
    return np.ascontiguousarray(planes), np.zeros((n, 48), np.float32)


# ---- tests/test_density_control.mojo ---------------------------------------

def test_pruning_transparent(oracle):
    planes, sh = _scene([(0, 0, 0), (1, 1, 1)], [(0.1,) * 3] * 2, [0.8, 0.001])
    cfg = oracle.train_config(prune_opacity_threshold=0.005)
    op, osh, origin, (cl, sp, pr) = oracle.densify_and_prune(planes, sh, np.zeros(2), np.zeros(2),
                                                             cfg, 10.0)
    assert pr == 1 and op.shape[1] == 1
    assert op[10, 0] == np.float32(0.8)


def test_cloning_under_reconstructed(oracle):
    planes, sh = _scene([(0, 0, 0)], [(0.005,) * 3], [0.8])
    cfg = oracle.train_config(densify_grad_threshold=0.0002)
    op, _, origin, (cl, sp, pr) = oracle.densify_and_prune(planes, sh, [0.01], [1], cfg, 10.0)
    assert cl == 1 and op.shape[1] == 2
    # the original, then the copy offset by 0.01 * scale
    np.testing.assert_array_equal(op[0:3, 0], [0, 0, 0])
    np.testing.assert_array_equal(op[0:3, 1], np.float32(0.01) * np.float32(0.005))
    assert origin.tolist() == [0, 1]


def test_splitting_over_reconstructed(oracle):
    planes, sh = _scene([(0, 0, 0)], [(0.2,) * 3], [0.8])
    cfg = oracle.train_config(densify_grad_threshold=0.0002)
    op, _, origin, (cl, sp, pr) = oracle.densify_and_prune(planes, sh, [0.01], [1], cfg, 10.0)
    assert sp == 1 and op.shape[1] == 2
    assert abs(op[3, 0] - 0.125) < 1e-4
    ns = np.float32(0.2) / np.float32(1.6)
    np.testing.assert_array_equal(op[0:3, 0], ns * np.float32(0.5))
    np.testing.assert_array_equal(op[0:3, 1], -(ns * np.float32(0.5)))
    assert origin.tolist() == [1, 1]


def test_opacity_reset(oracle):
    op = np.array([0.95, 0.005, 0.01, 0.5], np.float32)
    oracle.reset_opacities(op, 0.01)
    assert abs(op[0] - 0.01) < 1e-5
    np.testing.assert_array_equal(op, np.array([0.01, 0.005, 0.01, 0.01], np.float32))


# ---- invariants on a random scene -------------------------------------------

def test_densify_invariants(oracle):
    sc = random_scene(3000, 1.0, 81)
    planes = sc.planes.copy()
    rng = np.random.default_rng(3)
    planes[3:6] *= rng.choice([1.0, 30.0, 400.0], size=sc.count)  # some large, some huge
    planes[6:10] *= 3.0                                            # unnormalized rotations
    vs = rng.exponential(0.0003, sc.count).astype(np.float32)
    cnt = rng.integers(0, 4, sc.count).astype(np.int32)
    cfg = oracle.train_config()
    extent = 4.0
    op, osh, origin, (cl, sp, pr) = oracle.densify_and_prune(planes, sc.sh, vs, cnt, cfg, extent)
    src = origin // 2
    fresh = origin % 2 == 1
    n_out = op.shape[1]
    assert n_out == sc.count - pr + cl + sp
    assert np.all(np.diff(src) >= 0)                       # source order
    np.testing.assert_array_equal(osh, sc.sh[src])         # SH copied as is
    # who was pruned: low opacity, or a scale beyond half the extent
    max_s = planes[3:6].max(axis=0)
    pruned = (planes[10] < np.float32(0.005)) | (max_s > np.float32(extent) * np.float32(0.5))
    assert pr == int(pruned.sum()) and not np.any(pruned[src])
    # rotations are unit quaternions after the Gaussian3D normalization
    np.testing.assert_allclose(np.linalg.norm(op[6:10], axis=0), 1.0, atol=1e-6)
    # untouched Gaussians keep position, scale and opacity exactly
    keep = ~fresh
    np.testing.assert_array_equal(op[0:6, keep], planes[0:6, src[keep]])
    np.testing.assert_array_equal(op[10], planes[10, src])
    assert cl > 0 and sp > 0 and pr > 0


# ---- Adam -------------------------------------------------------------------

def _textbook_adam(p, grads, lr_of_step, b1=0.9, b2=0.999, eps=1e-8, clamp=None):
    m = np.zeros_like(p, np.float64)
    v = np.zeros_like(p, np.float64)
    p = p.astype(np.float64).copy()
    for t, g in enumerate(grads, start=1):
        g = g.astype(np.float64)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p = p - lr_of_step(t) * mh / (np.sqrt(vh) + eps)
        if clamp:
            p = np.clip(p, *clamp)
    return p


@pytest.mark.parametrize("all_params", [False, True])
def test_adam_matches_textbook(oracle, all_params):
    rng = np.random.default_rng(11)
    n, steps = 400, 6
    sc = random_scene(n, 1.0, 82)
    planes, sh = sc.planes.copy(), sc.sh.copy()
    p0, s0 = planes.copy(), sh.copy()
    m = (np.zeros_like(planes), np.zeros_like(sh))
    v = (np.zeros_like(planes), np.zeros_like(sh))
    cfg = oracle.train_config(iterations=50)
    gp = [rng.normal(0, 1e-3, planes.shape).astype(np.float32) for _ in range(steps)]
    gs = [rng.normal(0, 1e-3, sh.shape).astype(np.float32) for _ in range(steps)]
    for t in range(1, steps + 1):
        oracle.adam_step(planes, sh, gp[t - 1], gs[t - 1], m, v, cfg, t, all_params)
    lr_pos = lambda t: 0.00016 * (0.0000016 / 0.00016) ** min(1.0, t / 50)  # noqa: E731
    for k in range(11):
        lr = {0: lr_pos, 1: lr_pos, 2: lr_pos}.get(k)
        if k == 10:
            lr = lambda t: 0.05  # noqa: E731
        elif k >= 3 and lr is None:
            if not all_params:
                np.testing.assert_array_equal(planes[k], p0[k])
                continue
            lr = (lambda t: 0.005) if k < 6 else (lambda t: 0.001)  # noqa: E731
        ref = _textbook_adam(p0[k], [g[k] for g in gp], lr,
                             clamp=(0.01, 0.99) if k == 10 else None)
        np.testing.assert_allclose(planes[k], ref, rtol=2e-5, atol=1e-7)
    if all_params:
        ref = _textbook_adam(s0, gs, lambda t: 0.0025)
        np.testing.assert_allclose(sh, ref, rtol=2e-5, atol=1e-7)
    else:
        np.testing.assert_array_equal(sh, s0)


def test_adam_first_step_closed_form(oracle):
    g = np.array([[1e-3, -2e-2, 5.0, -7e-4]] * 11, np.float32)
    planes = np.full((11, 4), 0.5, np.float32)
    sh = np.zeros((4, 48), np.float32)
    m = (np.zeros_like(planes), np.zeros_like(sh))
    v = (np.zeros_like(planes), np.zeros_like(sh))
    cfg = oracle.train_config(iterations=1000)
    oracle.adam_step(planes, sh, g, np.zeros_like(sh), m, v, cfg, 1)
    lr = 0.00016 * (0.01 ** (1 / 1000))
    want = 0.5 - lr * g[0] / (np.abs(g[0]) + 1e-8)
    np.testing.assert_allclose(planes[0], want, rtol=1e-6)
    np.testing.assert_allclose(planes[10], 0.5 - 0.05 * np.sign(g[0]), rtol=1e-5)


def test_position_lr_schedule(oracle):
    cfg = oracle.train_config(iterations=100)
    assert oracle.adam_schedule(cfg, 100).pos_lr == pytest.approx(0.0000016, rel=1e-5)
    assert oracle.adam_schedule(cfg, 500).pos_lr == pytest.approx(0.0000016, rel=1e-5)
    assert oracle.adam_schedule(cfg, 50).pos_lr == pytest.approx((0.00016 * 0.0000016) ** 0.5,
                                                                 rel=1e-5)
    s = oracle.adam_schedule(cfg, 1)
    assert s.inv_b1 == pytest.approx(10.0, rel=1e-5) and s.inv_b2 == pytest.approx(1000.0, rel=1e-3)
