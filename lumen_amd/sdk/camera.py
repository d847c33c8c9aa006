"""Camera and CameraTrajectory of the Python SDK (python/lumen/camera.py:9-129).

Same constructor, fields and methods (forward_vector, view_matrix_4x4,
orbit, interpolate).  Added for the GPU path: rotation_quaternion() gives the
camera-to-world rotation (w, x, y, z) that lumen_rasterize derives from
position / target / up (columns right, down, forward, +Z forward and +Y down
in view space, the convention of src/scene/camera.mojo:69-79), computed in
double from the float32-rounded vectors in the same order of operations as
the library does, so that both give the same float32 quaternion; and
to_pinhole() the lumen_amd PinholeCamera lumen_rasterize would render with.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
from typing import List, Tuple

import numpy as np

from ..camera import CameraConfig, PinholeCamera

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_tanf = _libm.tanf
_tanf.restype = ctypes.c_float
_tanf.argtypes = [ctypes.c_float]


class Camera:
    """Pinhole camera looking from `position` at `target` (python/lumen/camera.py:9-28)."""

    def __init__(self, position: Tuple[float, float, float] = (0.0, 0.0, 3.0),
                 target: Tuple[float, float, float] = (0.0, 0.0, 0.0),
                 up: Tuple[float, float, float] = (0.0, 1.0, 0.0),
                 fov_x_deg: float = 60.0, fov_y_deg: float = 45.0,
                 near_plane: float = 0.1, far_plane: float = 100.0):
        self.position = list(position)
        self.target = list(target)
        self.up = list(up)
        self.fov_x_deg = float(fov_x_deg)
        self.fov_y_deg = float(fov_y_deg)
        self.near_plane = float(near_plane)
        self.far_plane = float(far_plane)

    def forward_vector(self) -> Tuple[float, float, float]:
        """Unit vector from position to target; (0, 0, -1) when they coincide."""
        d = [t - p for t, p in zip(self.target, self.position)]
        n = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        if n < 1e-6:
            return (0.0, 0.0, -1.0)
        return (d[0] / n, d[1] / n, d[2] / n)

    def view_matrix_4x4(self) -> List[List[float]]:
        """World-to-view look-at matrix with rows right, true up and -forward
        (python/lumen/camera.py:40-69)."""
        f = self.forward_vector()
        u0 = self.up
        r = [f[1] * u0[2] - f[2] * u0[1], f[2] * u0[0] - f[0] * u0[2], f[0] * u0[1] - f[1] * u0[0]]
        rn = math.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
        r = [1.0, 0.0, 0.0] if rn < 1e-6 else [v / rn for v in r]
        u = [r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]]
        p = self.position
        dot = lambda a: a[0] * p[0] + a[1] * p[1] + a[2] * p[2]  # noqa: E731
        return [[r[0], r[1], r[2], -dot(r)],
                [u[0], u[1], u[2], -dot(u)],
                [-f[0], -f[1], -f[2], dot(f)],
                [0.0, 0.0, 0.0, 1.0]]

    def rotation_quaternion(self) -> np.ndarray:
        """Camera-to-world (w, x, y, z) float32, as lumen_rasterize derives it."""
        f32 = lambda v: [float(np.float32(x)) for x in v]  # noqa: E731
        eye, tgt, up = f32(self.position), f32(self.target), f32(self.up)
        f = [tgt[k] - eye[k] for k in range(3)]
        fl = math.sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2])
        if fl < 1e-12:
            fl = 1.0
        f = [v / fl for v in f]
        r = [f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]]
        rl = math.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
        if rl < 1e-12:  # up along the view direction: any perpendicular axis
            a = [1.0, 0.0, 0.0] if abs(f[0]) < 0.9 else [0.0, 1.0, 0.0]
            r = [f[1] * a[2] - f[2] * a[1], f[2] * a[0] - f[0] * a[2], f[0] * a[1] - f[1] * a[0]]
            rl = math.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
        r = [v / rl for v in r]
        d = [f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]]
        m00, m01, m02 = r[0], d[0], f[0]
        m10, m11, m12 = r[1], d[1], f[1]
        m20, m21, m22 = r[2], d[2], f[2]
        tr = m00 + m11 + m22
        if tr > 0.0:
            s = 2.0 * math.sqrt(tr + 1.0)
            q = (0.25 * s, (m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s)
        elif m00 > m11 and m00 > m22:
            s = 2.0 * math.sqrt(1.0 + m00 - m11 - m22)
            q = ((m21 - m12) / s, 0.25 * s, (m01 + m10) / s, (m02 + m20) / s)
        elif m11 > m22:
            s = 2.0 * math.sqrt(1.0 + m11 - m00 - m22)
            q = ((m02 - m20) / s, (m01 + m10) / s, 0.25 * s, (m12 + m21) / s)
        else:
            s = 2.0 * math.sqrt(1.0 + m22 - m00 - m11)
            q = ((m10 - m01) / s, (m02 + m20) / s, (m12 + m21) / s, 0.25 * s)
        return np.array(q, np.float32)

    def focal_length(self, height: int) -> float:
        """H / (2 tanf(fov_y / 2)) in float32, lumen_rasterize's focal length
        (the C library's tanf, as the library calls it)."""
        f32 = np.float32
        fov = f32(f32(f32(self.fov_y_deg) * f32(math.pi)) / f32(180.0))
        half = f32(fov * f32(0.5))
        return float(f32(f32(height) / (f32(2.0) * f32(_tanf(float(half))))))

    def to_pinhole(self, width: int, height: int, near_plane: float,
                   far_plane: float) -> PinholeCamera:
        """The camera lumen_rasterize renders a width x height frame with: focal
        length from fov_y for both axes, principal point at the image centre,
        the given clip planes."""
        f = self.focal_length(height)
        cfg = CameraConfig(width, height, f, f, width * 0.5, height * 0.5, near_plane, far_plane)
        return PinholeCamera(cfg, np.array(self.position, np.float32), self.rotation_quaternion())


class CameraTrajectory:
    """Camera paths (python/lumen/camera.py:72-129)."""

    @staticmethod
    def orbit(target: Tuple[float, float, float] = (0.0, 0.0, 0.0), radius: float = 3.0,
              elevation_deg: float = 15.0, num_frames: int = 60,
              fov_y_deg: float = 45.0) -> List[Camera]:
        """num_frames cameras on a circle at the given elevation, azimuth
        2 pi i / num_frames measured from +Z towards +X, all facing the target."""
        el = math.radians(elevation_deg)
        y = target[1] + radius * math.sin(el)
        ring = radius * math.cos(el)
        cams = []
        for i in range(num_frames):
            az = 2.0 * math.pi * float(i) / float(num_frames)
            pos = (target[0] + ring * math.sin(az), y, target[2] + ring * math.cos(az))
            cams.append(Camera(position=pos, target=target, up=(0.0, 1.0, 0.0),
                               fov_y_deg=fov_y_deg))
        return cams

    @staticmethod
    def interpolate(cam_a: Camera, cam_b: Camera, t: float) -> Camera:
        """Linear blend of position, target and fov_y with t clamped to [0, 1];
        up and clip planes from cam_a."""
        t = max(0.0, min(1.0, t))
        s = 1.0 - t
        mix = lambda a, b: tuple(a[k] * s + b[k] * t for k in range(3))  # noqa: E731
        return Camera(position=mix(cam_a.position, cam_b.position),
                      target=mix(cam_a.target, cam_b.target), up=cam_a.up,
                      fov_y_deg=cam_a.fov_y_deg * s + cam_b.fov_y_deg * t,
                      near_plane=cam_a.near_plane, far_plane=cam_a.far_plane)
