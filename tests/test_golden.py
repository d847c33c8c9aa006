"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle).

CPU: the oracle reproduces the fixtures from their own inputs, and the scene
generator still produces the inputs the digests were taken from.
GPU: the HIP path reproduces them bit for bit (binning stages, images).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from lumen_amd.camera import CameraConfig, PinholeCamera
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import GaussianScene, workload

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def lattice():
    z = np.load(os.path.join(HERE, "lattice512_800x600.npz"))
    c = z["camera"]
    cam = PinholeCamera(CameraConfig(int(c[0]), int(c[1]), float(c[2]), float(c[3]), float(c[4]),
                                     float(c[5]), float(c[6]), float(c[7])),
                        c[8:11].astype(np.float32), c[11:15].astype(np.float32))
    return z, GaussianScene(z["planes"].copy(), z["sh"].copy()), cam


@pytest.fixture(scope="module")
def r100k():
    with open(os.path.join(HERE, "random100k_1280x720.json")) as f:
        return json.load(f)


def check_lattice_images(z, rgb, depth, t_final, n_eval):
    pix = z["pixels"]
    np.testing.assert_array_equal(rgb[pix[:, 0], pix[:, 1]], z["rgb_at"])
    np.testing.assert_array_equal(depth[pix[:, 0], pix[:, 1]], z["depth_at"])
    np.testing.assert_array_equal(t_final[pix[:, 0], pix[:, 1]], z["t_at"])
    np.testing.assert_array_equal(n_eval[pix[:, 0], pix[:, 1]], z["n_eval_at"])
    assert [sha(rgb), sha(depth), sha(t_final), sha(n_eval)] == list(z["digests"])


def test_oracle_reproduces_lattice_fixture(lattice, oracle):
    z, scene, cam = lattice
    ref = oracle.rasterize(scene, cam, RenderConfig())
    assert [ref["visible"], ref["instances"], ref["tiles"], ref["evaluations"]] == \
        list(z["counts"])
    for k in ("vis_to_orig", "radius", "rect", "keys_unsorted", "vals_unsorted", "keys", "vals",
              "ranges"):
        np.testing.assert_array_equal(ref[k], z[k], err_msg=k)
    check_lattice_images(z, ref["rgb"], ref["depth"], ref["t_final"], ref["n_eval"])


def test_lattice_fixture_is_sorted_and_partitioned(lattice):
    """Size-independent properties of the frozen binning stages."""
    z, _, _ = lattice
    keys, vals, ranges = z["keys"], z["vals"], z["ranges"]
    assert np.all(keys[1:] >= keys[:-1])
    order = np.lexsort((np.arange(keys.size), z["keys_unsorted"]))
    np.testing.assert_array_equal(keys, z["keys_unsorted"][order])
    np.testing.assert_array_equal(vals, z["vals_unsorted"][order])  # stable
    nonempty = ranges[ranges[:, 0] != -1]
    assert nonempty[0, 0] == 0 and nonempty[-1, 1] == keys.size
    assert np.all(nonempty[1:, 0] == nonempty[:-1, 1])


def test_scene_generator_matches_fixture_inputs(r100k):
    scene, _ = workload("random100k_1280x720")
    assert sha(scene.planes) == r100k["inputs_sha256"]["planes"]
    assert sha(scene.sh) == r100k["inputs_sha256"]["sh"]


@pytest.mark.gpu
def test_hip_reproduces_lattice_fixture(lattice, rasterizer):
    z, scene, cam = lattice
    out = rasterizer.render(scene, cam, RenderConfig(), keep_stages=True, save_pixels=True)
    assert out.visible_splats == z["counts"][0]
    assert out.total_splat_instances == z["counts"][1]
    np.testing.assert_array_equal(rasterizer.stage("vis_to_orig"), z["vis_to_orig"])
    np.testing.assert_array_equal(rasterizer.stage("radius"), z["radius"])
    for k in ("keys_unsorted", "vals_unsorted", "keys", "vals", "ranges"):
        np.testing.assert_array_equal(rasterizer.stage(k), z[k], err_msg=k)
    H, W = cam.config.height, cam.config.width
    check_lattice_images(z, out.rgb.cpu().numpy(), out.depth.cpu().numpy(),
                         rasterizer.stage("t_final").reshape(H, W),
                         rasterizer.stage("n_eval").reshape(H, W))


@pytest.mark.gpu
def test_hip_reproduces_random100k_digests(r100k, rasterizer):
    scene, cam = workload("random100k_1280x720")
    out = rasterizer.render(scene, cam, RenderConfig(), keep_stages=True)
    assert out.visible_splats == r100k["visible"]
    assert out.total_splat_instances == r100k["instances"]
    d = r100k["sha256"]
    assert sha(rasterizer.stage("keys")) == d["keys"]
    assert sha(rasterizer.stage("vals").astype(np.int64)) == d["vals"]
    assert sha(rasterizer.stage("ranges").astype(np.int64)) == d["ranges"]
    rgb, depth = out.rgb.cpu().numpy(), out.depth.cpu().numpy()
    pix = np.array(r100k["pixels"])
    np.testing.assert_array_equal(rgb[pix[:, 0], pix[:, 1]].astype(np.float64), r100k["rgb_at"])
    np.testing.assert_array_equal(depth[pix[:, 0], pix[:, 1]].astype(np.float64),
                                  r100k["depth_at"])
    assert sha(rgb) == d["rgb"] and sha(depth) == d["depth"]
