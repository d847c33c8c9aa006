"""Gaussian scenes in the GaussianSoA layout and the synthetic workloads.

GaussianSoA (src/scene/gaussian.mojo:47-107) is 11 FP32 planes of N values
(pos x y z, scale x y z, rot w x y z, opacity -- activated, not a logit) plus
an N x 48 SH block indexed g*48 + k*3 + c.  Host scenes are numpy; device
scenes hold one (11, N) tensor and one (N, 48) tensor so that every plane is a
contiguous, 16-byte aligned run the kernels read coalesced.

Workload generators (fixed seeds, numpy PCG64, float64 draws cast to FP32):
  lattice_scene         512 Gaussians on an 8^3 lattice over [-1, 1]^3,
                        jitter U(-0.02, 0.02), log-scales U(ln .03, ln .08),
                        random unit quaternions, opacity sigmoid(N(0, 1)),
                        SH DC N(0.5, 0.2), higher orders N(0, 0.05)
  random_scene          N Gaussians, means N(0, sigma^2 I), log-scales
                        N(ln .01, 0.5), random unit quaternions, opacity
                        sigmoid(N(0, 1.5)), SH as above
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

PLANE_NAMES = ("pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z",
               "rot_w", "rot_x", "rot_y", "rot_z", "opacity")
SH_COEFFS = 48


@dataclass
class GaussianScene:
    planes: np.ndarray  # (11, N) float32
    sh: np.ndarray      # (N, 48) float32

    def __post_init__(self):
        self.planes = np.ascontiguousarray(self.planes, np.float32)
        self.sh = np.ascontiguousarray(self.sh, np.float32)
        assert self.planes.ndim == 2 and self.planes.shape[0] == 11
        assert self.sh.shape == (self.planes.shape[1], SH_COEFFS)

    @property
    def count(self) -> int:
        return int(self.planes.shape[1])

    def subset(self, idx) -> "GaussianScene":
        idx = np.asarray(idx)
        return GaussianScene(self.planes[:, idx], self.sh[idx])

    @staticmethod
    def from_arrays(pos, scale, rot, opacity, sh) -> "GaussianScene":
        pos = np.asarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        planes = np.concatenate([pos.T, np.asarray(scale, np.float32).reshape(n, 3).T,
                                 np.asarray(rot, np.float32).reshape(n, 4).T,
                                 np.asarray(opacity, np.float32).reshape(1, n)], axis=0)
        return GaussianScene(planes, np.asarray(sh, np.float32).reshape(n, SH_COEFFS))

    def to_device(self, device="cuda") -> "DeviceScene":
        import torch
        return DeviceScene(torch.from_numpy(self.planes).to(device),
                           torch.from_numpy(self.sh).to(device))


class DeviceScene:
    """A scene resident in device memory (torch tensors used as plain buffers)."""

    def __init__(self, planes, sh):
        import torch
        assert planes.dtype == torch.float32 and sh.dtype == torch.float32
        self.planes = planes.contiguous()
        self.sh = sh.contiguous()
        assert self.planes.shape[0] == 11 and self.sh.shape == (self.planes.shape[1], SH_COEFFS)

    @property
    def count(self) -> int:
        return int(self.planes.shape[1])

    @property
    def device(self):
        return self.planes.device

    def nbytes(self) -> int:
        return self.planes.numel() * 4 + self.sh.numel() * 4

    def ex_struct(self):
        from ._lib import LumenExScene
        s = LumenExScene()
        n = self.count
        s.count = n
        base = self.planes.data_ptr()
        for k, name in enumerate(PLANE_NAMES):
            setattr(s, name, base + 4 * k * n)
        s.sh = self.sh.data_ptr()
        return s


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def _unit_quaternions(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def _sh(rng, n):
    sh = rng.normal(0.0, 0.05, (n, SH_COEFFS)).astype(np.float32)
    sh[:, 0:3] = rng.normal(0.5, 0.2, (n, 3))
    return sh


def lattice_scene(seed: int = 1001) -> GaussianScene:
    rng = np.random.default_rng(seed)
    coords = -1.0 + 2.0 * np.arange(8) / 7.0
    X, Y, Z = np.meshgrid(coords, coords, coords, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    n = pos.shape[0]
    pos = pos + rng.uniform(-0.02, 0.02, (n, 3))
    scale = np.exp(rng.uniform(math.log(0.03), math.log(0.08), (n, 3)))
    rot = _unit_quaternions(rng, n)
    opacity = _sigmoid(rng.standard_normal(n))
    return GaussianScene.from_arrays(pos, scale, rot, opacity, _sh(rng, n))


def random_scene(n: int, mean_sigma: float = 1.0, seed: int = 1002,
                 chunk: int = 1 << 20) -> GaussianScene:
    """Generated in chunks of `chunk` Gaussians so that multi-million scenes
    stay within a few hundred MB of float64 temporaries."""
    rng = np.random.default_rng(seed)
    planes = np.empty((11, n), np.float32)
    sh = np.empty((n, SH_COEFFS), np.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = b - a
        planes[0:3, a:b] = (rng.standard_normal((m, 3)) * mean_sigma).T
        planes[3:6, a:b] = np.exp(rng.normal(math.log(0.01), 0.5, (m, 3))).T
        planes[6:10, a:b] = _unit_quaternions(rng, m).T
        planes[10, a:b] = _sigmoid(rng.normal(0.0, 1.5, m))
        sh[a:b] = _sh(rng, m)
    return GaussianScene(planes, sh)


def perturbed(scene: GaussianScene, sigma: float = 0.01, seed: int = 7) -> GaussianScene:
    """Copy with means + N(0, sigma^2) (the target scene of the fwd+bwd workload)."""
    rng = np.random.default_rng(seed)
    planes = scene.planes.copy()
    planes[0:3] += rng.normal(0.0, sigma, (3, scene.count)).astype(np.float32)
    return GaussianScene(planes, scene.sh.copy())


# ---- named workloads --------------------------------------------------------

def workload(name: str):
    """(scene, camera) of a named workload."""
    from .camera import CameraConfig, look_at
    if name == "lattice512_800x600":
        cfg = CameraConfig(800, 600, 700.0, 700.0, 400.0, 300.0, 0.01, 1000.0)
        return lattice_scene(1001), look_at(cfg, (0.0, 0.0, 4.0))
    if name == "random100k_1280x720":
        cfg = CameraConfig(1280, 720, 1000.0, 1000.0, 640.0, 360.0, 0.01, 1000.0)
        return random_scene(100_000, 1.0, 1002), look_at(cfg, (0.0, 0.0, 5.0))
    if name == "random1m_1920x1080":
        cfg = CameraConfig(1920, 1080, 1400.0, 1400.0, 960.0, 540.0, 0.01, 1000.0)
        return random_scene(1_000_000, math.sqrt(2.0), 1003), look_at(cfg, (0.0, 0.0, 5.0))
    raise KeyError(name)


def orbit_workload(n: int = 3_000_000, views: int = 32, radius: float = 6.0):
    """The multi-view workload: n Gaussians (the 1M workload's generator,
    means N(0, 2 I)), `views` 1920x1080 cameras on a closed Catmull-Rom orbit
    of the given radius looking at the origin."""
    from .camera import CameraConfig, orbit
    cfg = CameraConfig(1920, 1080, 1400.0, 1400.0, 960.0, 540.0, 0.01, 1000.0)
    return random_scene(n, math.sqrt(2.0), 1004), orbit(cfg, views, radius)


SERVING_SCENES = (10_000, 100_000, 500_000, 1_000_000, 2_000_000, 3_000_000, 3_000_000,
                  5_000_000)
SERVING_RESOLUTIONS = ((320, 240, 233.3), (1280, 720, 933.3), (1920, 1080, 1400.0),
                       (3840, 2160, 2800.0))


def serving_scenes(sizes=SERVING_SCENES):
    """The serving workload's resident scenes (same generator, seeds 1100+k),
    named s<k>_<n>."""
    return {f"s{k}_{n}": random_scene(n, math.sqrt(2.0), 1100 + k) for k, n in enumerate(sizes)}
