"""Writes the golden fixtures under tests/golden/ from the oracle.

    python tests/golden/make_golden.py

lattice512_800x600.npz  the 512-Gaussian lattice scene (inputs included, so the
                        fixture does not depend on the scene generator), its
                        camera, every binning stage (visible map, radii, keys
                        and values before and after the sort, tile ranges)
                        and 512 sampled pixels of rgb / depth / T / evaluations,
                        plus SHA-256 digests of the full images.
random100k_1280x720.json  counts and SHA-256 digests of the sorted keys,
                        values, ranges, rgb and depth of the 100k-Gaussian
                        workload (inputs regenerated from its seed; their
                        digest is stored too), plus sampled pixels.

The oracle is pinned separately (reference known-answer tests, closed forms,
brute force); these files freeze its outputs so a regression on either side
shows without rerunning it.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lumen_amd.raster import RenderConfig  # noqa: E402
from lumen_amd.scene import workload  # noqa: E402
from oracle import pyoracle  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_pixels(H, W, count, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, H, count), rng.integers(0, W, count)], axis=1).astype(np.int32)


def camera_array(cam):
    c = cam.config
    return np.array([c.width, c.height, c.focal_x, c.focal_y, c.principal_x, c.principal_y,
                     c.near_plane, c.far_plane, *cam.position, *cam.rotation], np.float64)


def lattice():
    scene, cam = workload("lattice512_800x600")
    ref = pyoracle.rasterize(scene, cam, RenderConfig())
    H, W = ref["rgb"].shape[:2]
    pix = sample_pixels(H, W, 512, 5)
    np.savez_compressed(
        os.path.join(HERE, "lattice512_800x600.npz"),
        planes=scene.planes, sh=scene.sh, camera=camera_array(cam),
        counts=np.array([ref["visible"], ref["instances"], ref["tiles"], ref["evaluations"]],
                        np.int64),
        vis_to_orig=ref["vis_to_orig"], radius=ref["radius"], rect=ref["rect"],
        keys_unsorted=ref["keys_unsorted"], vals_unsorted=ref["vals_unsorted"],
        keys=ref["keys"], vals=ref["vals"], ranges=ref["ranges"], pixels=pix,
        rgb_at=ref["rgb"][pix[:, 0], pix[:, 1]], depth_at=ref["depth"][pix[:, 0], pix[:, 1]],
        t_at=ref["t_final"][pix[:, 0], pix[:, 1]], n_eval_at=ref["n_eval"][pix[:, 0], pix[:, 1]],
        digests=np.array([sha(ref["rgb"]), sha(ref["depth"]), sha(ref["t_final"]),
                          sha(ref["n_eval"])]))
    print("lattice512:", ref["visible"], ref["instances"], ref["evaluations"])


def random100k():
    scene, cam = workload("random100k_1280x720")
    ref = pyoracle.rasterize(scene, cam, RenderConfig())
    H, W = ref["rgb"].shape[:2]
    pix = sample_pixels(H, W, 64, 6)
    out = {
        "workload": "random100k_1280x720",
        "inputs_sha256": {"planes": sha(scene.planes), "sh": sha(scene.sh)},
        "visible": int(ref["visible"]), "instances": int(ref["instances"]),
        "tiles": int(ref["tiles"]), "evaluations": int(ref["evaluations"]),
        "sha256": {"keys": sha(ref["keys"]), "vals": sha(ref["vals"].astype(np.int64)),
                   "ranges": sha(ref["ranges"].astype(np.int64)), "rgb": sha(ref["rgb"]),
                   "depth": sha(ref["depth"])},
        "pixels": pix.tolist(),
        "rgb_at": ref["rgb"][pix[:, 0], pix[:, 1]].astype(np.float64).tolist(),
        "depth_at": ref["depth"][pix[:, 0], pix[:, 1]].astype(np.float64).tolist(),
    }
    with open(os.path.join(HERE, "random100k_1280x720.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("random100k:", out["visible"], out["instances"], out["evaluations"])


if __name__ == "__main__":
    lattice()
    random100k()
