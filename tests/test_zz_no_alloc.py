"""apt_trace_closest_host_noalloc (the entry point that allocates nothing)
through the C ABI: bit-identical hits to apt_trace_closest_host on two calls
in a row, the second with fewer rays, and sizes beyond its fixed device array
refused."""
import ctypes as C

import numpy as np
import pytest

from appearance_amd._ffi import lib
from appearance_amd.raytrace import HIT_DTYPE, RAY_DTYPE, Bvh
from tests.test_oracle_bvh import random_rays, random_tris

pytestmark = pytest.mark.gpu
MAX_RAYS = 65536
APT_ERR_INVALID_ARG = 1


def _noalloc(bvh, rays):
    L = lib()
    fn = L.apt_trace_closest_host_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    rays = np.ascontiguousarray(rays).view(RAY_DTYPE).reshape(-1)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    rc = fn(bvh.handle, rays.ctypes.data, rays.shape[0], hits.ctypes.data)
    return rc, hits


def test_noalloc_matches_allocating_entry(gpu_available):
    tris = random_tris(30000, 11)
    bvh = Bvh(tris, width=4)
    for n, seed in ((MAX_RAYS, 12), (1000, 13)):
        print(f"[noalloc] {n} rays", flush=True)
        rays = np.ascontiguousarray(random_rays(n, seed)).view(RAY_DTYPE).reshape(-1)
        rc, got = _noalloc(bvh, rays)
        assert rc == 0, lib().apt_last_error()
        want = bvh.trace_closest(rays)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        assert (got["prim"] != 0xFFFFFFFF).sum() > n // 10


def test_noalloc_limits(gpu_available):
    tris = random_tris(2000, 14)
    bvh = Bvh(tris, width=4)
    rays = np.ascontiguousarray(random_rays(MAX_RAYS + 1, 15)).view(RAY_DTYPE).reshape(-1)
    rc, _ = _noalloc(bvh, rays)
    assert rc == APT_ERR_INVALID_ARG
    assert b"65536" in lib().apt_last_error()
    rc, hits = _noalloc(bvh, rays[:0])
    assert rc == 0 and hits.shape == (0,)
