"""Image export and terminal preview of the Python SDK
(python/lumen/visualizer.py:8-82), vectorised over the pixels."""
from __future__ import annotations

from typing import Sequence

import numpy as np


def export_ppm(filepath: str, width: int, height: int, rgb_bytes: bytes) -> None:
    """Binary PPM (P6) of an 8-bit RGB buffer."""
    with open(filepath, "wb") as f:
        f.write(f"P6\n{width} {height}\n255\n".encode("ascii"))
        f.write(rgb_bytes)


def export_depth_visualization(filepath: str, width: int, height: int,
                               depth_values: Sequence[float], near_plane: float = 0.5,
                               far_plane: float = 10.0) -> None:
    """Depth as a false-colour PPM: t = clamp((d - near) / (far - near), 0, 1)
    (t = 1 where d >= 999 or d <= 0), i = int(255 (1 - t)), colour
    (i, int(0.8 i), int(255 - 0.7 i)) clamped to [0, 255]."""
    d = np.asarray(depth_values[: width * height], np.float64)
    empty = (d >= 999.0) | (d <= 0.0)
    t = np.clip((d - near_plane) / (far_plane - near_plane), 0.0, 1.0)
    t[empty] = 1.0
    i = ((1.0 - t) * 255.0).astype(np.int64)
    rgb = np.stack([i, (i * 0.8).astype(np.int64), (255 - i * 0.7).astype(np.int64)], axis=1)
    export_ppm(filepath, width, height, np.clip(rgb, 0, 255).astype(np.uint8).tobytes())


def ascii_render(width: int, height: int, alpha_values: Sequence[float], target_width: int = 60,
                 target_height: int = 25) -> str:
    """Nearest-sample preview of the alpha plane with the ramp ' .:-=+*#%@'."""
    ramp = np.array(list(" .:-=+*#%@"))
    a = np.asarray(alpha_values[: width * height], np.float64).reshape(height, width)
    ys = np.minimum((np.arange(target_height) * (height / target_height)).astype(np.int64), height - 1)
    xs = np.minimum((np.arange(target_width) * (width / target_width)).astype(np.int64), width - 1)
    levels = (np.clip(a[np.ix_(ys, xs)], 0.0, 1.0) * (len(ramp) - 1)).astype(np.int64)
    return "\n".join("".join(row) for row in ramp[levels])
