"""Ray-query level API: BVH build and closest/any-hit traversal on the GPU.

Mirrors what the reference obtains from the Vulkan driver: the BLAS build
(src/scene_resources/scene_model.rs:53-88) and the ray query
(rayQueryInitialize/Proceed/GetCommittedIntersection, trace.wgsl:159-162;
terminate-on-first-hit shadow queries, helpers/trace.wgsl:20-26).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import BvhOpts, BvhStats, TlasStats, check, lib, vptr

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4")])
# apt_instance (64 B): object-to-world 3x4 row major, BLAS index, custom index
INSTANCE_DTYPE = np.dtype([("transform", "<f4", (3, 4)), ("blas", "<u4"), ("custom_index", "<u4"),
                           ("_pad", "<u4", 2)])
assert INSTANCE_DTYPE.itemsize == 64
MAX_INSTANCES = 8192
# bytes of one wide node (quantized child boxes, include/apt.h apt_bvh_export)
NODE_BYTES = {2: 48, 4: 64, 8: 96}

BUILDER_HOST = 0
BUILDER_GPU = 1


def plan(tris: np.ndarray, width: int = 4, max_leaf: int = 4, bins: int = 32):
    """Host build without the device (apt_bvh_plan): (status, statistics dict).
    A nonzero status means the tree needs a deeper traversal stack than
    stack_capacity() entries and every tracing entry point refuses it."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    st = BvhStats()
    opts = BvhOpts(width, max_leaf, bins, BUILDER_HOST, 8, 0)
    rc = lib().apt_bvh_plan(_ffi.ptr(tris), tris.shape[0], C.byref(opts), C.byref(st))
    return rc, {f: getattr(st, f) for f, _ in BvhStats._fields_}


def stack_capacity() -> int:
    return int(lib().apt_trace_stack_capacity())


class Bvh:
    """A wide BVH resident in device memory (libapt's apt_bvh)."""

    def __init__(self, tris: np.ndarray, width: int = 4, max_leaf: int = 4, bins: int = 32,
                 builder: int = BUILDER_HOST, lds_stack: int = 8, refill: int = 0):
        tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        self.n_tris = tris.shape[0]
        opts = BvhOpts(width, max_leaf, bins, builder, lds_stack, refill)
        h = C.c_void_p()
        check(lib().apt_bvh_build(_ffi.ptr(tris), self.n_tris, C.byref(opts), C.byref(h)),
              "apt_bvh_build")
        self._h = h

    @classmethod
    def from_device(cls, d_tri, n_tris: int, width: int = 4, max_leaf: int = 4, bins: int = 32,
                    builder: int = BUILDER_GPU, lds_stack: int = 8, refill: int = 0) -> "Bvh":
        """Build from n_tris * 9 floats already in device memory (apt_bvh_build_device);
        d_tri is a device pointer or an object with data_ptr() (a torch tensor)."""
        ptr = d_tri.data_ptr() if hasattr(d_tri, "data_ptr") else int(d_tri)
        self = cls.__new__(cls)
        self.n_tris = int(n_tris)
        opts = BvhOpts(width, max_leaf, bins, builder, lds_stack, refill)
        h = C.c_void_p()
        check(lib().apt_bvh_build_device(C.c_void_p(ptr), self.n_tris, C.byref(opts), C.byref(h)),
              "apt_bvh_build_device")
        self._h = h
        return self

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def stats(self) -> dict:
        s = BvhStats()
        check(lib().apt_bvh_get_stats(self._h, C.byref(s)), "apt_bvh_get_stats")
        return {f: getattr(s, f) for f, _ in BvhStats._fields_}

    def export(self):
        """(node bytes, triangle records (n_tris*3, 4) float32) of the device layout."""
        st = self.stats()
        nodes = np.empty(st["n_nodes"] * NODE_BYTES[st["width"]], np.uint8)
        tris = np.empty((st["n_tris"] * 3, 4), np.float32)
        check(lib().apt_bvh_export(self._h, vptr(nodes), nodes.size, vptr(tris), tris.nbytes),
              "apt_bvh_export")
        return nodes, tris

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        """Host arrays in and out (synchronous)."""
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE).reshape(-1)
        hits = np.empty(rays.shape[0], dtype=HIT_DTYPE)
        check(lib().apt_trace_closest_host(self._h, vptr(rays), rays.shape[0], vptr(hits)),
              "apt_trace_closest_host")
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE).reshape(-1)
        occ = np.empty(rays.shape[0], dtype=np.uint32)
        check(lib().apt_trace_any_host(self._h, vptr(rays), rays.shape[0], vptr(occ)),
              "apt_trace_any_host")
        return occ

    def trace_closest_device(self, d_rays, n: int, d_hits, stream=None) -> None:
        """Device pointers (torch tensors or raw ints); asynchronous on stream."""
        check(lib().apt_trace_closest(self._h, _dp(d_rays), n, _dp(d_hits), _dp(stream)),
              "apt_trace_closest")

    def trace_any_device(self, d_rays, n: int, d_occ, stream=None) -> None:
        check(lib().apt_trace_any(self._h, _dp(d_rays), n, _dp(d_occ), _dp(stream)),
              "apt_trace_any")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().apt_bvh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tlas:
    """A two-level scene (libapt's apt_tlas): BLASes copied into one device
    arena and a TLAS over instances of them, the reference's per-mesh BLAS and
    per-frame TLAS (src/scene_resources/scene_model.rs:53-88,
    scene_resources/mod.rs:159-276).  Hits carry custom_index + BLAS prim."""

    def __init__(self, width: int = 4, lds_stack: int = 8, refill: int = 0):
        opts = BvhOpts(width, 0, 0, 0, lds_stack, refill)
        h = C.c_void_p()
        check(lib().apt_tlas_create(C.byref(opts), C.byref(h)), "apt_tlas_create")
        self._h = h
        self.width = width

    def add_blas(self, bvh: Bvh) -> int:
        idx = C.c_uint32()
        check(lib().apt_tlas_add_blas(self._h, bvh.handle, C.byref(idx)), "apt_tlas_add_blas")
        return int(idx.value)

    def set_instances(self, inst: np.ndarray) -> None:
        inst = np.ascontiguousarray(inst, dtype=INSTANCE_DTYPE).reshape(-1)
        check(lib().apt_tlas_set_instances(self._h, vptr(inst), inst.shape[0]), "apt_tlas_set_instances")

    def stats(self) -> dict:
        s = TlasStats()
        check(lib().apt_tlas_get_stats(self._h, C.byref(s)), "apt_tlas_get_stats")
        return {f: getattr(s, f) for f, _ in TlasStats._fields_}

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE).reshape(-1)
        hits = np.empty(rays.shape[0], dtype=HIT_DTYPE)
        check(lib().apt_trace_closest_tlas_host(self._h, vptr(rays), rays.shape[0], vptr(hits)),
              "apt_trace_closest_tlas_host")
        return hits

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays).view(RAY_DTYPE).reshape(-1)
        occ = np.empty(rays.shape[0], dtype=np.uint32)
        check(lib().apt_trace_any_tlas_host(self._h, vptr(rays), rays.shape[0], vptr(occ)),
              "apt_trace_any_tlas_host")
        return occ

    def trace_closest_device(self, d_rays, n: int, d_hits, stream=None) -> None:
        check(lib().apt_trace_closest_tlas(self._h, _dp(d_rays), n, _dp(d_hits), _dp(stream)),
              "apt_trace_closest_tlas")

    def trace_any_device(self, d_rays, n: int, d_occ, stream=None) -> None:
        check(lib().apt_trace_any_tlas(self._h, _dp(d_rays), n, _dp(d_occ), _dp(stream)),
              "apt_trace_any_tlas")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().apt_tlas_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def instances(transforms, blas, custom) -> np.ndarray:
    """An apt_instance array from (n, 3, 4) transforms, BLAS indices and custom indices."""
    t = np.asarray(transforms, np.float32).reshape(-1, 3, 4)
    out = np.zeros(t.shape[0], INSTANCE_DTYPE)
    out["transform"] = t
    out["blas"] = blas
    out["custom_index"] = custom
    return out


def _dp(x):
    if x is None:
        return None
    if isinstance(x, int):
        return C.c_void_p(x)
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    if hasattr(x, "cuda_stream"):
        return C.c_void_p(x.cuda_stream)
    return x


def rays_bvh_stress(seed: int = 2, side: int = 1024, n_incoherent: int = 1 << 20) -> np.ndarray:
    """The BVH-stress ray sets: side*side coherent + n_incoherent random rays."""
    out = np.empty(side * side + n_incoherent, dtype=RAY_DTYPE)
    check(lib().apt_rays_bvh_stress(seed, side, n_incoherent, vptr(out)), "apt_rays_bvh_stress")
    return out
