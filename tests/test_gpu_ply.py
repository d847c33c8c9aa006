"""Scenes loaded from PLY files render on the GPU exactly as the oracle does.

- The 512-Gaussian lattice written in the Mojo core's ASCII layout (columns
  x y z f_dc0-2 opacity scale0-2 rot0-3, src/scene/ply.mojo:32-138) and read
  back with its parse_float arithmetic: every stage bit-exact against the
  oracle rendering the same loaded scene.
- A 100k-Gaussian scene through the binary loader used for resident serving
  scenes (exact round trip): the loaded scene renders bit for bit as the
  scene it was written from.
"""
import sys

import numpy as np
import pytest

from lumen_amd import ply
from lumen_amd.raster import RenderConfig
from lumen_amd.scene import random_scene, workload

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_lattice_from_mojo_ascii_ply(rasterizer, oracle, tmp_path):
    import torch
    sc, cam = workload("lattice512_800x600")
    path = str(tmp_path / "lattice512.ply")
    ply.write_ply_ascii(sc, path)
    loaded = ply.read_ply_ascii(path)
    assert loaded.count == 512
    # the decimal text round trip moves values by at most a few ulp
    np.testing.assert_allclose(loaded.planes, sc.planes, rtol=1e-5, atol=1e-6)
    say("[ply] lattice512 from the Mojo ASCII layout")
    cfg = RenderConfig()
    out = rasterizer.render(loaded.to_device(), cam, cfg, keep_stages=True, save_pixels=True)
    torch.cuda.synchronize()
    ref = oracle.rasterize(loaded, cam, cfg)
    assert out.visible_splats == ref["visible"] == 512
    assert out.total_splat_instances == ref["instances"]
    np.testing.assert_array_equal(rasterizer.stage("vis_to_orig"), ref["vis_to_orig"])
    np.testing.assert_array_equal(rasterizer.stage("keys"), ref["keys"])
    np.testing.assert_array_equal(rasterizer.stage("vals"), ref["vals"])
    np.testing.assert_array_equal(rasterizer.stage("ranges"), ref["ranges"])
    np.testing.assert_array_equal(out.rgb.cpu().numpy(), ref["rgb"])
    np.testing.assert_array_equal(out.depth.cpu().numpy(), ref["depth"])


def test_binary_ply_scene_renders_as_the_original(rasterizer, tmp_path):
    import torch
    sc, cam = workload("random100k_1280x720")
    path = str(tmp_path / "s100k.ply")
    ply.write_ply_binary(sc, path)
    loaded = ply.read_ply(path)
    say("[ply] 100k scene through the binary loader")
    cfg = RenderConfig()
    a = rasterizer.render(sc.to_device(), cam, cfg)
    rgb_a, depth_a = a.rgb.clone(), a.depth.clone()
    b = rasterizer.render(loaded.to_device(), cam, cfg)
    torch.cuda.synchronize()
    assert a.total_splat_instances == b.total_splat_instances
    assert torch.equal(rgb_a, b.rgb) and torch.equal(depth_a, b.depth)
