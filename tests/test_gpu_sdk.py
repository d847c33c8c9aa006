"""The Python SDK (lumen_amd.sdk) rendering on the GPU.

python/tests/test_client.py:83-140 (render a cube, stats, centre alpha,
ASCII preview, image export) restated against LumenEngine, whose render
goes through lumen_rasterize.  The frame is then pinned to the oracle, and
render_batch (one multi-view lumen_ex_render_batch frame over an orbit) must
return, per camera, exactly what render returns."""
import os
import sys

import numpy as np
import pytest

from lumen_amd.raster import RenderConfig
from lumen_amd.sdk import (Camera, CameraTrajectory, GaussianCloud, LumenEngine, RasterizerConfig,
                           ascii_render, export_depth_visualization, export_ppm)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_rasterize_scene():
    """python/tests/test_client.py:84-117."""
    cfg = RasterizerConfig(image_width=80, image_height=60, tile_size=16, near_plane=0.1,
                           far_plane=50.0)
    engine = LumenEngine(cfg)
    cube = GaussianCloud.synthetic_cube(points_per_axis=3, cube_size=1.0)
    res = engine.render(cube, Camera(position=(0.0, 0.0, 2.5), target=(0.0, 0.0, 0.0)))
    say(f"[sdk] stats {res.stats}")
    assert (res.width, res.height) == (80, 60)
    assert len(res.rgb_bytes) == 80 * 60 * 3
    assert len(res.depth_values) == 80 * 60 and len(res.alpha_values) == 80 * 60
    assert res.stats["total_splats"] == 27
    assert res.stats["visible_splats"] > 0 and res.stats["touched_tiles"] > 0
    assert res.alpha_values[(60 // 2) * 80 + (80 // 2)] > 0.0
    art = ascii_render(res.width, res.height, res.alpha_values, 30, 15)
    assert isinstance(art, str) and len(art) > 0
    engine.close()


def test_image_export(tmp_path):
    """python/tests/test_client.py:119-140."""
    engine = LumenEngine(RasterizerConfig(image_width=40, image_height=30))
    res = engine.render(GaussianCloud.synthetic_cube(points_per_axis=2, cube_size=0.8),
                        Camera(position=(0.0, 0.0, 2.0)))
    ppm, dep = str(tmp_path / "render.ppm"), str(tmp_path / "depth.ppm")
    export_ppm(ppm, res.width, res.height, res.rgb_bytes)
    export_depth_visualization(dep, res.width, res.height, res.depth_values)
    assert os.path.getsize(ppm) > 0 and os.path.getsize(dep) > 0
    engine.close()


@pytest.mark.parametrize("eye", [(0.0, 0.0, 3.0), (1.2, 0.8, 2.4)])
def test_engine_frame_matches_oracle(oracle, eye):
    W, H = 160, 120
    cfg = RasterizerConfig(image_width=W, image_height=H, near_plane=0.1, far_plane=50.0)
    engine = LumenEngine(cfg)
    cloud = GaussianCloud.synthetic_sphere(300, 1.0)
    cam = Camera(position=eye)
    res = engine.render(cloud, cam)
    ref = oracle.rasterize(cloud.to_scene(), cam.to_pinhole(W, H, 0.1, 50.0), RenderConfig())
    assert res.stats["visible_splats"] == ref["visible"]
    rgb8 = np.frombuffer(res.rgb_bytes, np.uint8).reshape(H, W, 3).astype(int)
    exp8 = (np.minimum(ref["rgb"], np.float32(1.0)) * np.float32(255.0)).astype(np.uint8)
    diff = np.abs(rgb8 - exp8.astype(int))
    say(f"[sdk] eye {eye}: max |rgb8 - oracle| = {diff.max()}")
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
    a_ref = 1.0 - ref["t_final"].reshape(-1)
    np.testing.assert_allclose(np.array(res.alpha_values), np.where(a_ref > 0, a_ref, 0.0),
                               atol=1e-4)
    engine.close()


def test_render_batch_equals_render():
    W, H = 160, 120
    engine = LumenEngine(RasterizerConfig(image_width=W, image_height=H))
    cloud = GaussianCloud.synthetic_sphere(400, 1.0)
    cams = CameraTrajectory.orbit(radius=3.0, num_frames=6)
    batch = engine.render_batch(cloud, cams)
    assert len(batch) == len(cams)
    for k, cam in enumerate(cams):
        one = engine.render(cloud, cam)
        assert batch[k].rgb_bytes == one.rgb_bytes, k
        assert batch[k].alpha_values == one.alpha_values, k
        assert batch[k].depth_values == one.depth_values, k
        for key in ("total_splats", "visible_splats", "culled_splats", "total_tiles",
                    "touched_tiles"):
            assert batch[k].stats[key] == one.stats[key], (k, key)
    say(f"[sdk] render_batch of {len(cams)} orbit views equals render per view")
    engine.close()
