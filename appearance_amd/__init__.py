"""appearance_amd: Appearance's path-tracing hot path, native to MI355X (gfx950).

The product is libapt.so (include/apt.h): HIP kernels for BVH traversal and the
wavefront integrator plus a C++ host runtime.  This package is the Python view
of that C ABI (ctypes), used by tests and bench.py.
"""
from ._ffi import AptError, lib  # noqa: F401
from .raytrace import Bvh, RAY_DTYPE, HIT_DTYPE, rays_bvh_stress  # noqa: F401
from .scene import SceneData, generate_scene  # noqa: F401
