"""The reference's Python SDK surface (python/lumen/__init__.py) on liblumen:
the same names, fields and call signatures, rendering on the GPU."""
from .camera import Camera, CameraTrajectory
from .client import LumenEngine, RasterizerConfig, RenderResult
from .cloud import Gaussian3D, GaussianCloud
from .visualizer import ascii_render, export_depth_visualization, export_ppm

__version__ = "0.1.0"
__all__ = ["Camera", "CameraTrajectory", "GaussianCloud", "Gaussian3D", "LumenEngine",
           "RasterizerConfig", "RenderResult", "export_ppm", "export_depth_visualization",
           "ascii_render"]
