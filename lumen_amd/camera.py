"""Cameras: CameraConfig / PinholeCamera mirrors and a look-at that faces its target.

CameraConfig follows src/core/config.mojo:17-42 (same field names and
defaults).  PinholeCamera holds (config, position, rotation) like
src/scene/camera.mojo:24-40; the rotation is camera-to-world (w, x, y, z), the
camera looks along +Z with +Y down in view space (camera.mojo:69-79).

The reference's PinholeCamera.look_at (camera.mojo:43-64) builds an OpenGL
(-Z forward) rotation and drops to identity when the matrix trace is <= 0, so
it faces away from its target.  look_at here builds R = [right | down |
forward] with forward = normalize(target - eye), right = normalize(forward x
up), down = forward x right, and converts R to a quaternion with Shepperd's
method (all four branches).  Both the oracle and the GPU path normalize the
quaternion again on use, so parity does not depend on its rounding.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import List, Sequence

import numpy as np


@dataclass
class CameraConfig:
    width: int = 1280
    height: int = 720
    focal_x: float = 800.0
    focal_y: float = 800.0
    principal_x: float = 640.0
    principal_y: float = 360.0
    near_plane: float = 0.01
    far_plane: float = 1000.0


@dataclass
class PinholeCamera:
    config: CameraConfig
    position: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 5.0], np.float32))
    rotation: np.ndarray = field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0], np.float32))

    def __post_init__(self):
        self.position = np.asarray(self.position, np.float32).reshape(3)
        self.rotation = np.asarray(self.rotation, np.float32).reshape(4)

    def with_resolution(self, width: int, height: int, focal: float) -> "PinholeCamera":
        cfg = replace(self.config, width=width, height=height, focal_x=focal, focal_y=focal,
                      principal_x=width / 2.0, principal_y=height / 2.0)
        return PinholeCamera(cfg, self.position.copy(), self.rotation.copy())


def rotation_to_quaternion(R: np.ndarray) -> np.ndarray:
    """Shepperd's method; R is a proper rotation (float64).  Returns (w, x, y, z)."""
    m = np.asarray(R, np.float64)
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0.0:
        s = 2.0 * math.sqrt(tr + 1.0)
        q = (0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s)
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = 2.0 * math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2])
        q = ((m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s)
    elif m[1, 1] > m[2, 2]:
        s = 2.0 * math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2])
        q = ((m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s)
    else:
        s = 2.0 * math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1])
        q = ((m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s)
    return np.array(q, np.float64)


def quaternion_to_rotation(q: Sequence[float]) -> np.ndarray:
    w, x, y, z = (float(v) for v in q)
    n = math.sqrt(w * w + x * x + y * y + z * z)
    w, x, y, z = w / n, x / n, y / n, z / n
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def look_at(config: CameraConfig, eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> PinholeCamera:
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    if np.linalg.norm(r) < 1e-12:
        r = np.cross(f, np.array([1.0, 0.0, 0.0]) if abs(f[0]) < 0.9 else np.array([0.0, 1.0, 0.0]))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f], axis=1)
    q = rotation_to_quaternion(R)
    return PinholeCamera(config, eye.astype(np.float32), q.astype(np.float32))


def catmull_rom(p0, p1, p2, p3, t: float) -> np.ndarray:
    """catmull_rom_spline, src/scene/trajectory.mojo:24-42 (float64 here)."""
    t2 = t * t
    t3 = t2 * t
    f1 = -0.5 * t3 + t2 - 0.5 * t
    f2 = 1.5 * t3 - 2.5 * t2 + 1.0
    f3 = -1.5 * t3 + 2.0 * t2 + 0.5 * t
    f4 = 0.5 * t3 - 0.5 * t2
    return (np.asarray(p0) * f1 + np.asarray(p1) * f2 + np.asarray(p2) * f3 + np.asarray(p3) * f4)


def orbit(config: CameraConfig, n_views: int, radius: float, height: float = 1.0,
          n_keyframes: int = 8, target=(0.0, 0.0, 0.0)) -> List[PinholeCamera]:
    """n_views cameras along a closed Catmull-Rom spline through n_keyframes
    points on a circle of the given radius, each looking at the target."""
    keys = [np.array([radius * math.cos(2 * math.pi * k / n_keyframes), height,
                      radius * math.sin(2 * math.pi * k / n_keyframes)]) for k in range(n_keyframes)]
    cams = []
    for v in range(n_views):
        s = v * n_keyframes / n_views
        k = int(math.floor(s))
        t = s - k
        p = catmull_rom(keys[(k - 1) % n_keyframes], keys[k % n_keyframes],
                        keys[(k + 1) % n_keyframes], keys[(k + 2) % n_keyframes], t)
        cams.append(look_at(config, p, target))
    return cams
