"""gprt_amd — MI355X (gfx950) software acceleration-structure path of GPRT.

The product is libgprt_hip.so (C-ABI declared in include/gprt_hip.h, HIP sources in csrc/).
This package loads it through ctypes (api.py) and provides seeded scene / ray generators
(scenes.py). There is no Python or CPU fallback for any entry point.
"""
from . import api  # noqa: F401
from .api import Buffer, Context, TriangleAccel, load  # noqa: F401

__all__ = ["api", "scenes", "Buffer", "Context", "TriangleAccel", "load"]
