"""Gaussian3D and GaussianCloud of the Python SDK (python/lumen/cloud.py:11-199).

Same fields, constructors, generators and PLY methods (save_ply / load_ply in
the SDK's own column order: x y z scale_0..2 rot_0..3 opacity f_dc_0..2,
%.6f, python/lumen/cloud.py:133-199).  Added: aos() / to_scene() /
from_scene() convert to the layouts the GPU path takes (the 236-byte
LumenGaussian records of lumen.h and the GaussianSoA planes), and
save_ply_core / load_ply_core use the Mojo core's column order
(lumen_amd.ply).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from ..scene import GaussianScene

_C0 = 0.2820947917
_SDK_COLUMNS = ("x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2",
                "rot_3", "opacity", "f_dc_0", "f_dc_1", "f_dc_2")


@dataclass
class Gaussian3D:
    position: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    scale: Tuple[float, float, float] = (0.05, 0.05, 0.05)
    rotation: Tuple[float, float, float, float] = (1.0, 0.0, 0.0, 0.0)  # (w, x, y, z)
    opacity: float = 0.8
    sh_coeffs: List[float] = field(default_factory=lambda: [0.0] * 48)


class GaussianCloud:
    def __init__(self, splats: Optional[List[Gaussian3D]] = None):
        self.splats: List[Gaussian3D] = splats if splats is not None else []

    def __len__(self) -> int:
        return len(self.splats)

    def add(self, splat: Gaussian3D) -> None:
        self.splats.append(splat)

    def bounding_box(self):
        """((min x, y, z), (max x, y, z)) of the centres; zeros when empty."""
        if not self.splats:
            return ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
        cols = list(zip(*(g.position for g in self.splats)))
        return (tuple(min(c) for c in cols), tuple(max(c) for c in cols))

    def centroid(self) -> Tuple[float, float, float]:
        if not self.splats:
            return (0.0, 0.0, 0.0)
        n = float(len(self.splats))
        cols = list(zip(*(g.position for g in self.splats)))
        return tuple(sum(c) / n for c in cols)

    @classmethod
    def synthetic_cube(cls, points_per_axis: int = 4, cube_size: float = 1.0) -> "GaussianCloud":
        """points_per_axis^3 splats on a lattice over [-size/2, size/2]^3 (x
        outermost), colour from the normalised position through the DC term,
        isotropic scale 0.6 * spacing (0.1 for a single point), opacity 0.85."""
        step = cube_size / float(points_per_axis - 1) if points_per_axis > 1 else 0.0
        half = cube_size * 0.5
        s = step * 0.6 if step > 0 else 0.1
        cloud = cls()
        for xi in range(points_per_axis):
            for yi in range(points_per_axis):
                for zi in range(points_per_axis):
                    p = (-half + xi * step, -half + yi * step, -half + zi * step)
                    sh = [0.0] * 48
                    for c in range(3):
                        norm = (p[c] + half) / cube_size if cube_size > 0 else 0.5
                        sh[c] = (norm - 0.5) / _C0
                    cloud.add(Gaussian3D(p, (s, s, s), (1.0, 0.0, 0.0, 0.0), 0.85, sh))
        return cloud

    @classmethod
    def synthetic_sphere(cls, point_count: int = 128, radius: float = 1.0) -> "GaussianCloud":
        """Fibonacci points on a sphere, colour from the normalised position,
        isotropic scale 0.12 * radius, opacity 0.9."""
        golden = 1.6180339887
        cloud = cls()
        s = radius * 0.12
        for i in range(point_count):
            theta = 2.0 * math.pi * float(i) / golden
            phi = math.acos(1.0 - 2.0 * (float(i) + 0.5) / float(point_count))
            p = (radius * math.sin(phi) * math.cos(theta), radius * math.sin(phi) * math.sin(theta),
                 radius * math.cos(phi))
            sh = [0.0] * 48
            for c in range(3):
                sh[c] = (0.5 + 0.5 * (p[c] / radius) - 0.5) / _C0
            cloud.add(Gaussian3D(p, (s, s, s), (1.0, 0.0, 0.0, 0.0), 0.90, sh))
        return cloud

    # ---- layouts of the GPU path --------------------------------------------

    def aos(self) -> np.ndarray:
        """(N, 59) float32 rows in the LumenGaussian layout of lumen.h
        (position, scale, rotation w x y z, opacity, 48 SH coefficients)."""
        out = np.zeros((len(self.splats), 59), np.float32)
        for i, g in enumerate(self.splats):
            out[i, 0:3] = g.position
            out[i, 3:6] = g.scale
            out[i, 6:10] = g.rotation
            out[i, 10] = g.opacity
            sh = g.sh_coeffs[:48]
            out[i, 11:11 + len(sh)] = sh
        return out

    def to_scene(self) -> GaussianScene:
        a = self.aos()
        return GaussianScene(np.ascontiguousarray(a[:, :11].T), np.ascontiguousarray(a[:, 11:]))

    @classmethod
    def from_scene(cls, scene: GaussianScene) -> "GaussianCloud":
        p = scene.planes.astype(np.float64)
        sh = scene.sh.astype(np.float64)
        return cls([Gaussian3D(tuple(p[0:3, i]), tuple(p[3:6, i]), tuple(p[6:10, i]),
                               float(p[10, i]), list(sh[i])) for i in range(scene.count)])

    # ---- PLY ------------------------------------------------------------------

    def save_ply(self, filepath: str) -> None:
        """ASCII PLY in the SDK's column order, six decimals."""
        with open(filepath, "w", encoding="utf-8") as f:
            f.write(f"ply\nformat ascii 1.0\nelement vertex {len(self.splats)}\n")
            for k in _SDK_COLUMNS:
                f.write(f"property float {k}\n")
            f.write("end_header\n")
            for g in self.splats:
                vals = list(g.position) + list(g.scale) + list(g.rotation) + [g.opacity] + \
                    list(g.sh_coeffs[:3])
                f.write(" ".join(f"{v:.6f}" for v in vals) + "\n")

    @classmethod
    def load_ply(cls, filepath: str) -> "GaussianCloud":
        """Reads lines after end_header with at least 14 tokens in the SDK's
        column order; the higher SH coefficients are zero."""
        cloud = cls()
        in_header = True
        with open(filepath, "r", encoding="utf-8") as f:
            for line in f:
                s = line.strip()
                if in_header:
                    in_header = s != "end_header"
                    continue
                parts = s.split()
                if len(parts) < 14:
                    continue
                v = [float(x) for x in parts[:14]]
                sh = [0.0] * 48
                sh[0:3] = v[11:14]
                cloud.add(Gaussian3D(tuple(v[0:3]), tuple(v[3:6]), tuple(v[6:10]), v[10], sh))
        return cloud

    def save_ply_core(self, filepath: str) -> None:
        """ASCII PLY in the Mojo core's column order (write_ply_ascii)."""
        from ..ply import write_ply_ascii
        write_ply_ascii(self.to_scene(), filepath)

    @classmethod
    def load_ply_core(cls, filepath: str) -> "GaussianCloud":
        """read_ply_ascii of the Mojo core."""
        from ..ply import read_ply_ascii
        return cls.from_scene(read_ply_ascii(filepath))
