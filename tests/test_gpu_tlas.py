"""Two-level scenes on the GPU (apt_tlas_*) against the oracle's instance
traversal (tests/test_oracle_instances.py pins the oracle).

Bar: bit-exact t, u, v, prim and any-hit flags, every BVH width, both
traversal schedules (persistent pool and one ray per thread)."""
import numpy as np
import pytest

from appearance_amd._ffi import AptError
from appearance_amd.raytrace import BUILDER_GPU, INSTANCE_DTYPE, Bvh, Tlas
from tests._oracle import hits_as_struct
from tests.instances_util import (dyadic_rays, dyadic_tris, exact_transform, general_transform,
                                  make_instances, world_soup)
from tests.test_oracle_bvh import random_rays, random_tris

pytestmark = pytest.mark.gpu


def _check(got, ref):
    gt, gu, gv, gp = got["t"], got["u"], got["v"], got["prim"]
    rt, ru, rv, rp = hits_as_struct(ref)
    np.testing.assert_array_equal(gp, rp)
    h = rp != 0xFFFFFFFF
    np.testing.assert_array_equal(gt[h], rt[h])
    np.testing.assert_array_equal(gu[h], ru[h])
    np.testing.assert_array_equal(gv[h], rv[h])
    assert np.all(np.isinf(gt[~h]))


def _scene(blas_tris, inst, width, lds_stack=8, refill=0, builder=0):
    t = Tlas(width=width, lds_stack=lds_stack, refill=refill)
    for b in blas_tris:
        t.add_blas(Bvh(b, width=width, builder=builder))
    t.set_instances(inst)
    return t


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("refill", [0, 65])
def test_exact_instances_bit_exact(oracle, gpu_available, width, refill):
    print(f"[tlas] exact instances W={width} refill={refill}", flush=True)
    rng = np.random.default_rng(5)
    blas = [dyadic_tris(500, 51), dyadic_tris(200, 52, spread=2.0, size=0.8), dyadic_tris(1, 53)]
    xf = [np.hstack([np.eye(3), np.zeros((3, 1))])] + [exact_transform(rng) for _ in range(23)]
    ids = [k % 3 for k in range(24)]
    inst = make_instances(xf, ids, [b.shape[0] for b in blas])
    rays = dyadic_rays(4000, 54)
    t = _scene(blas, inst, width, refill=refill)
    ref = oracle.instances(blas).trace(inst, rays, brute=True)
    _check(t.trace_closest(rays), ref)
    # and the world soup traced as one BVH gives the same bits
    _check(t.trace_closest(rays), oracle.brute(world_soup(blas, inst), rays))
    rays[:, 7] = np.random.default_rng(6).uniform(1.0, 20.0, rays.shape[0])
    np.testing.assert_array_equal(t.trace_any(rays), oracle.instances(blas).trace(inst, rays, True, brute=True))


@pytest.mark.parametrize("width", [2, 4, 8])
def test_general_instances_bit_exact(oracle, gpu_available, width):
    print(f"[tlas] general instances W={width}", flush=True)
    rng = np.random.default_rng(9)
    blas = [random_tris(3000, 91), random_tris(800, 92, spread=1.5, size=0.3)]
    xf = [general_transform(rng) for _ in range(40)]
    for _ in range(24):  # small instances far from the ray origins
        m = general_transform(rng, scale=(2e-3, 6e-3), shift=1.0)
        m[:, 3] = rng.uniform(-40, 40, 3)
        xf.append(m)
    ids = [k % 2 for k in range(len(xf))]
    inst = make_instances(xf, ids, [b.shape[0] for b in blas])
    rays = random_rays(20000, 93, spread=30.0)
    t = _scene(blas, inst, width, builder=BUILDER_GPU)
    ref = oracle.instances(blas).trace(inst, rays)
    _check(t.trace_closest(rays), ref)
    assert (ref[:, 3].view(np.uint32) != 0xFFFFFFFF).sum() > 4000
    rays[:, 7] = 12.0
    np.testing.assert_array_equal(t.trace_any(rays), oracle.instances(blas).trace(inst, rays, True))


def test_many_instances_and_update(oracle, gpu_available):
    """1024 instances of 4 BLASes, 65,536 rays; then new transforms on the same scene."""
    print("[tlas] 1024 instances + update", flush=True)
    rng = np.random.default_rng(13)
    blas = [random_tris(2000, 130 + k, spread=0.6, size=0.08) for k in range(4)]
    sizes = [b.shape[0] for b in blas]

    def grid_instances(jitter):
        xf = []
        for k in range(1024):
            m = general_transform(rng, scale=(0.5, 1.5), shift=0.0)
            i, j, l = k % 16, (k // 16) % 8, k // 128
            m[:, 3] = np.array([i - 7.5, j - 3.5, l - 3.5]) * 1.3 + rng.uniform(-jitter, jitter, 3)
            xf.append(m)
        return make_instances(xf, [k % 4 for k in range(1024)], sizes)

    t = _scene(blas, grid_instances(0.1), 4, builder=BUILDER_GPU)
    rays = random_rays(65536, 131, spread=14.0)
    for jitter in (0.1, 0.3):
        inst = grid_instances(jitter)
        t.set_instances(inst)
        st = t.stats()
        assert st["n_instances"] == 1024 and st["n_blas"] == 4
        ref = oracle.instances(blas).trace(inst, rays)
        _check(t.trace_closest(rays), ref)
        assert (ref[:, 3].view(np.uint32) != 0xFFFFFFFF).sum() > 10000


def test_identity_instance_equals_plain_bvh(gpu_available):
    tris = random_tris(5000, 21)
    rays = random_rays(8000, 22)
    bvh = Bvh(tris, width=4)
    plain = bvh.trace_closest(rays.view(np.float32).reshape(-1, 8))
    t = Tlas(width=4)
    t.add_blas(bvh)
    inst = np.zeros(1, INSTANCE_DTYPE)
    inst["transform"] = np.hstack([np.eye(3), np.zeros((3, 1))])
    t.set_instances(inst)
    got = t.trace_closest(rays)
    for f in ("t", "u", "v", "prim"):
        np.testing.assert_array_equal(got[f], plain[f])


def test_empty_and_bad_input(gpu_available):
    t = Tlas(width=4)
    b = t.add_blas(Bvh(random_tris(100, 31), width=4))
    rays = random_rays(300, 32)
    t.set_instances(np.zeros(0, INSTANCE_DTYPE))
    got = t.trace_closest(rays)
    assert np.all(got["prim"] == 0xFFFFFFFF) and np.all(np.isinf(got["t"]))
    assert not t.trace_any(rays).any()
    inst = np.zeros(1, INSTANCE_DTYPE)
    inst["blas"] = b  # zero matrix: singular
    with pytest.raises(AptError) as e:
        t.set_instances(inst)
    assert e.value.status == 4
    inst["transform"] = np.hstack([np.eye(3), np.zeros((3, 1))])
    inst["blas"] = 7
    with pytest.raises(AptError) as e:
        t.set_instances(inst)
    assert e.value.status == 1
    with pytest.raises(AptError):
        t.set_instances(np.zeros(8193, INSTANCE_DTYPE))
    with pytest.raises(AptError):
        t.add_blas(Bvh(random_tris(100, 33), width=8))
