"""The legacy C renderer's oracle and its exp, pinned on the CPU.

- orc_legacy_rasterize (oracle/lumen_oracle_legacy.c, a restatement of
  c_api/src/lumen_c.c:380-575) reproduces the fixtures that the reference's
  own C library wrote (tests/golden/legacy_c_renderer.npz, from
  tests/golden/make_legacy_golden.py): 8-bit RGB, depth and alpha planes and
  stats, bit for bit, over several frames and configs.
- Where that library is built here (oracle/_ref, from the reference's
  source), the restatement also equals it on random scenes.
- The C test's assertions (c_api/test/test_c_api.c:127-172) hold.
- legacy_expf (lumen_amd/csrc/legacy_expf.h, the expf the GPU's legacy mode
  uses) equals the C library's expf at every float x <= 0 (all 2^31 of
  them, NaNs included), and its table is 2^(i/32) rounded to double.
"""
import os
import struct
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "legacy_c_renderer.npz")
REF = os.path.join(ROOT, "oracle", "_ref", "liblumen_ref.so")
CASES = ("c_test_cube", "sphere128", "scene512", "scene512b", "away")


def load_case(name):
    z = np.load(GOLDEN)
    return {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}


def case_args(c):
    cam = c["camera"]
    camera = {"pos": cam[0:3], "target": cam[3:6], "up": cam[6:9], "fov_x": cam[9],
              "fov_y": cam[10], "near": cam[11], "far": cam[12]}
    cf = c["config"]
    config = {"near": cf[0], "far": cf[1], "eta": cf[2], "sigma": cf[3], "tile": int(c["tile"][0])}
    return camera, config


def stats_array(st):
    return np.array([st["total"], st["visible"], st["culled"], st["total_tiles"], st["touched"]],
                    np.uint64)


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_fixtures(oracle, name):
    c = load_case(name)
    camera, config = case_args(c)
    rgb = c["init_rgb"].copy()
    depth = c["init_depth"].copy()
    alpha = c["init_alpha"].copy()
    for k in range(int(c["frames"][0])):
        st = oracle.legacy_rasterize(c["aos"], camera, config, rgb, depth, alpha)
        np.testing.assert_array_equal(rgb, c[f"rgb{k}"])
        np.testing.assert_array_equal(depth.view(np.uint32), c[f"depth{k}"].view(np.uint32))
        np.testing.assert_array_equal(alpha.view(np.uint32), c[f"alpha{k}"].view(np.uint32))
        np.testing.assert_array_equal(stats_array(st), c[f"stats{k}"])


def test_fixtures_hold_the_c_test_contract():
    # c_api/test/test_c_api.c:160-171
    c = load_case("c_test_cube")
    st = c["stats0"]
    assert st[0] == 64 and st[1] > 0 and st[4] > 0
    assert c["alpha0"][60, 80] > 0.0
    # the exact depth ties of scene512 exercise the sort's tie order
    z = load_case("scene512")["aos"][:, 2]
    assert len(np.unique(z)) == 8


def random_scene_aos(n, seed, op_max=1.2):
    rng = np.random.default_rng(seed)
    aos = np.zeros((n, 59), np.float32)
    aos[:, 0:3] = rng.normal(0, 1.0, (n, 3))
    aos[:, 2] = np.round(aos[:, 2] * 4) / 4  # many exact depth ties
    aos[:, 3:6] = np.exp(rng.normal(np.log(0.05), 0.6, (n, 3)))
    aos[:, 6] = 1.0
    aos[:, 10] = rng.uniform(0.0, op_max, n)  # opacities above 1 too
    aos[:, 11:14] = rng.normal(0, 1.0, (n, 3))
    return aos


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("seed,tile,eta,sigma", [(1, 16, 0.99, 3.0), (2, 5, 0.9, 2.0),
                                                 (3, 64, 2.0, 3.5)])
def test_oracle_equals_reference_library(oracle, seed, tile, eta, sigma):
    """The restatement against the reference's C library itself, on random
    scenes with depth ties, opacities above 1 and (eta = 2) pixels whose alpha
    passes 1."""
    from tests.golden.make_legacy_golden import ref_lib, run_case
    L = ref_lib()
    aos = random_scene_aos(3000, seed, 3.0 if eta > 1.0 else 1.2)
    W, H = 197, 133
    cam = ((0.4, -0.3, 4.0), (0, 0, 0), (0, 1, 0), 60.0, 50.0, 0.1, 100.0)
    cfg = (tile, 0.2, 30.0, eta, sigma)
    ref = run_case(L, "r", aos, W, H, cam, cfg, frames=2, clear=(10, 200, 30))
    ref = {k.split("/", 1)[1]: v for k, v in ref.items()}
    camera, config = case_args(ref)
    rgb, depth, alpha = ref["init_rgb"].copy(), ref["init_depth"].copy(), ref["init_alpha"].copy()
    for k in range(2):
        st = oracle.legacy_rasterize(aos, camera, config, rgb, depth, alpha)
        np.testing.assert_array_equal(rgb, ref[f"rgb{k}"])
        np.testing.assert_array_equal(depth.view(np.uint32), ref[f"depth{k}"].view(np.uint32))
        np.testing.assert_array_equal(alpha.view(np.uint32), ref[f"alpha{k}"].view(np.uint32))
        np.testing.assert_array_equal(stats_array(st), ref[f"stats{k}"])
    if eta > 1.0:
        assert float(alpha.max()) > 1.0


def test_legacy_expf_table_is_2_pow_i_over_32():
    getcontext().prec = 60
    text = open(os.path.join(ROOT, "lumen_amd", "csrc", "legacy_expf.h")).read()
    body = text[text.index("#define LUMEN_LEGACY_EXP2_TABLE"):]
    body = body[body.index("{") + 1:body.index("}")]
    table = [int(t.strip().rstrip("ull"), 16) for t in body.replace("\\", " ").split(",")]
    assert len(table) == 32
    for i, t in enumerate(table):
        v = float(Decimal(2) ** (Decimal(i) / Decimal(32)))
        bits = struct.unpack("<Q", struct.pack("<d", v))[0]
        assert t == (bits - (i << 47)) % (1 << 64), i


HARNESS = r"""
#include <math.h>
#include <stdio.h>
#include "legacy_expf.h"
int main(void) {
  unsigned long long bad = 0, n = 0;
  for (unsigned long long b = 0x80000000ull; b <= 0xFFFFFFFFull; ++b) {
    uint32_t u = (uint32_t)b, ua, uc;
    float x, a, c;
    memcpy(&x, &u, 4);
    a = expf(x);
    c = legacy_expf(x);
    memcpy(&ua, &a, 4);
    memcpy(&uc, &c, 4);
    ++n;
    if (ua != uc && !(a != a && c != c)) ++bad;
  }
  const float pos[] = {0.0f, 1e-30f, 0.5f, 1.0f, 20.0f, 88.0f, 88.7f, 89.0f, INFINITY};
  for (int i = 0; i < 9; ++i) {
    float a = expf(pos[i]), c = legacy_expf(pos[i]);
    if (memcmp(&a, &c, 4) != 0) ++bad;
  }
  printf("%llu %llu\n", n, bad);
  return 0;
}
"""


def test_legacy_expf_equals_libm_expf_for_every_nonpositive_float(tmp_path):
    src = tmp_path / "h.c"
    src.write_text(HARNESS)
    exe = tmp_path / "h"
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "lumen_amd", "csrc"), str(src), "-o", str(exe), "-lm"],
                   check=True)
    n, bad = (int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                              check=True).stdout.split())
    assert n == 2 ** 31 and bad == 0
