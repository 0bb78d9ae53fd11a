"""Pins the oracle's two-level traversal (instances of BLASes).

The restatement (oracle_core.cpp, after scene_resources/mod.rs:159-276 and the
driver's ray query) maps each ray into the instance's object space with the
inverse of the instance transform and takes the lexicographic minimum of
(t, custom + BLAS prim) over all instances.  What pins it:

* the inverse: exact on translations, power-of-two scales, cyclic axis
  permutations and sign flips; W M = I to float rounding on random matrices;
  singular matrices refused;
* exact instances: for those transforms (and dyadic geometry) the instance
  query gives the same bits as the plain brute force on the world-space soup,
  closest and any hit, so a transposed inverse, a wrong translation sign, a
  wrong prim offset or a missed instance fails;
* general instances (rotation, non-uniform scale, translation) against an
  FP64 intersection of the world-space soup: same prim, hit point within
  2e-5 of the coordinates' magnitude;
* the per-BLAS BVH2 path equals the brute force bit for bit, including small
  instances far from the ray origins.

Planted mistakes in the oracle (a transposed world-to-object matrix, a negated
translation) fail the exact-instance pin.
"""
import ctypes as C

import numpy as np
import pytest

from tests._oracle import hits_as_struct
from tests.instances_util import (dyadic_rays, dyadic_tris, exact_transform, fp64_closest,
                                  general_transform, make_instances, world_soup)


class plant:
    def __init__(self, oracle, kind):
        self.o, self.k = oracle, kind

    def __enter__(self):
        self.o.lib.or_set_plant.argtypes = [C.c_int]
        self.o.lib.or_set_plant(self.k)

    def __exit__(self, *exc):
        self.o.lib.or_set_plant(0)


def _exact_scene(seed, n_inst=10):
    rng = np.random.default_rng(seed)
    blas = [dyadic_tris(300, seed + 1), dyadic_tris(120, seed + 2, spread=2.0, size=0.8),
            dyadic_tris(1, seed + 3)]
    xf = [np.hstack([np.eye(3), np.zeros((3, 1))])] + [exact_transform(rng) for _ in range(n_inst - 1)]
    ids = [int(rng.integers(len(blas))) for _ in range(n_inst)]
    inst = make_instances(xf, ids, [b.shape[0] for b in blas])
    return blas, inst


def test_affine_inverse_closed_forms(oracle):
    t = np.array([1.5, -2.25, 3.0], np.float32)
    m = np.hstack([np.eye(3), t[:, None]]).astype(np.float32)
    w = oracle.affine_inverse(m)
    np.testing.assert_array_equal(w, np.hstack([np.eye(3), -t[:, None]]).astype(np.float32))
    rng = np.random.default_rng(0)
    for _ in range(20):
        m = exact_transform(rng)
        w = oracle.affine_inverse(m.astype(np.float32))
        A = m[:, :3]
        want = np.hstack([np.linalg.inv(A), (-np.linalg.inv(A) @ m[:, 3])[:, None]])
        np.testing.assert_array_equal(w, want.astype(np.float32))  # exact for these
    for _ in range(50):
        m = general_transform(rng, scale=(0.05, 20.0)).astype(np.float32)
        w = oracle.affine_inverse(m).astype(np.float64)
        M = np.vstack([m.astype(np.float64), [0, 0, 0, 1]])
        W = np.vstack([w, [0, 0, 0, 1]])
        cond = np.linalg.cond(M[:3, :3])
        assert np.abs(W @ M - np.eye(4)).max() < 4e-7 * cond * 40
    assert oracle.affine_inverse(np.zeros((3, 4), np.float32)) is None
    sing = np.array([[1, 2, 3, 0], [2, 4, 6, 1], [0, 0, 1, 0]], np.float32)
    assert oracle.affine_inverse(sing) is None


@pytest.mark.parametrize("seed", [1, 2])
def test_exact_instances_equal_world_brute(oracle, seed):
    blas, inst = _exact_scene(10 * seed)
    soup = world_soup(blas, inst)
    rays = dyadic_rays(1500, 100 + seed)
    scene = oracle.instances(blas)
    ref = oracle.brute(soup, rays)
    for brute in (True, False):
        got = scene.trace(inst, rays, brute=brute)
        for a, b in zip(hits_as_struct(got), hits_as_struct(ref)):
            np.testing.assert_array_equal(a, b)
    hit = hits_as_struct(ref)[3] != 0xFFFFFFFF
    assert hit.sum() > 300
    # each instance alone, with rays aimed at it
    for k in range(len(inst)):
        one = inst[k:k + 1]
        tk = world_soup(blas, one)
        rk = dyadic_rays(200, 1000 * seed + k, center=tk.reshape(-1, 3).mean(0), aim=0.5)
        rk_ref = oracle.brute(tk, rk)
        rk_ref[:, 3] = (rk_ref[:, 3].view(np.uint32) + np.where(
            rk_ref[:, 3].view(np.uint32) == 0xFFFFFFFF, 0, one["custom_index"][0]).astype(np.uint32)).view(np.float32)
        got = scene.trace(one, rk, brute=True)
        for a, b in zip(hits_as_struct(got), hits_as_struct(rk_ref)):
            np.testing.assert_array_equal(a, b)
        assert (hits_as_struct(got)[3] != 0xFFFFFFFF).sum() > 0
    rays_any = rays.copy()
    rays_any[:, 7] = np.random.default_rng(seed).uniform(1.0, 20.0, rays.shape[0])
    ref_any = oracle.brute(soup, rays_any, True)
    for brute in (True, False):
        np.testing.assert_array_equal(scene.trace(inst, rays_any, any_hit=True, brute=brute), ref_any)


@pytest.mark.parametrize("kind", [7, 8])
def test_exact_instance_pin_catches_planted_mistakes(oracle, kind):
    blas, inst = _exact_scene(10)
    soup = world_soup(blas, inst)
    rays = dyadic_rays(1500, 101)
    ref = oracle.brute(soup, rays)
    with plant(oracle, kind):
        got = oracle.instances(blas).trace(inst, rays, brute=True)
    assert (hits_as_struct(got)[3] != hits_as_struct(ref)[3]).sum() > 50


def test_general_instances_vs_fp64(oracle):
    rng = np.random.default_rng(7)
    blas = [dyadic_tris(200, 71), dyadic_tris(80, 72, spread=1.5)]
    xf = [general_transform(rng) for _ in range(8)]
    ids = [k % 2 for k in range(8)]
    inst = make_instances(xf, ids, [b.shape[0] for b in blas])
    soup64 = world_soup(blas, inst, np.float64)
    rays = dyadic_rays(1200, 73)
    t64, p64, amb = fp64_closest(soup64, rays)
    got = oracle.instances(blas).trace(inst, rays, brute=True)
    gt, _, _, gp = hits_as_struct(got)
    clear = ~amb
    assert clear.mean() > 0.95
    np.testing.assert_array_equal(gp[clear], p64[clear])
    h = clear & (p64 != 0xFFFFFFFF)
    assert h.sum() > 300
    # the hit point agrees to 2e-5 of the coordinates' magnitude (float32
    # vertices and rays, grazing hits amplify the error in t)
    dn = np.linalg.norm(rays[:, 4:7].astype(np.float64), axis=1)
    scale = (np.linalg.norm(rays[:, 0:3].astype(np.float64), axis=1) + t64 * dn) / dn
    assert np.all(np.abs(gt[h] - t64[h]) <= 2e-5 * scale[h])
    assert np.median(np.abs(gt[h] - t64[h]) / t64[h]) < 1e-6


def test_instance_bvh_equals_instance_brute(oracle):
    rng = np.random.default_rng(11)
    blas = [dyadic_tris(400, 111), dyadic_tris(60, 112, spread=0.5, size=0.2)]
    xf = [general_transform(rng) for _ in range(12)]
    # small instances far from every ray origin: object-space origins lie
    # thousands of BLAS extents away
    for _ in range(6):
        m = general_transform(rng, scale=(2e-3, 6e-3), shift=1.0)
        m[:, 3] = rng.uniform(-40, 40, 3)
        xf.append(m)
    ids = [k % 2 for k in range(len(xf))]
    inst = make_instances(xf, ids, [b.shape[0] for b in blas])
    rays = dyadic_rays(1500, 113, spread=30.0)
    scene = oracle.instances(blas)
    a = scene.trace(inst, rays, brute=True)
    b = scene.trace(inst, rays, brute=False)
    for x, y in zip(hits_as_struct(a), hits_as_struct(b)):
        np.testing.assert_array_equal(x, y)
    rays[:, 7] = 25.0
    np.testing.assert_array_equal(scene.trace(inst, rays, any_hit=True, brute=True),
                                  scene.trace(inst, rays, any_hit=True, brute=False))
