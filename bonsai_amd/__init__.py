"""MI355X-native ray traversal and path integration for bonsai scenes.

The compute path is libbonsai_gpu.so (HIP for gfx950) behind the C ABI in
include/bonsai_gpu.h; this package is its Python host binding.
"""
from . import _native
from .gpu import HIT_DTYPE, Context, GpuBvh
from .scenes import Camera, SceneData

__all__ = ["Context", "GpuBvh", "Camera", "SceneData", "HIT_DTYPE", "_native"]
