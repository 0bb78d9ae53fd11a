"""Pins the oracle's canonical BVH2 traversal against brute force.

Closest hit = lexicographic min (t, prim) over all triangles; any hit =
existence.  Includes exact duplicate triangles (equal t, tie to lower prim)."""
import numpy as np

from tests._oracle import hits_as_struct


def random_tris(n, seed, spread=5.0, size=0.4):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 3))
    e1 = rng.normal(size=(n, 3)) * size
    e2 = rng.normal(size=(n, 3)) * size
    v0 = c
    return np.concatenate([v0, v0 + e1, v0 + e2], axis=1).astype(np.float32)


def random_rays(n, seed, spread=7.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = 1e30
    return r


def test_closest_matches_brute(oracle):
    tris = random_tris(3000, 1)
    tris = np.concatenate([tris, tris[:500]])  # duplicates: ties resolve to the lower prim
    rays = random_rays(4000, 2)
    b = oracle.bvh(tris)
    got = b.trace(rays)
    ref = oracle.brute(tris, rays)
    gt, gu, gv, gp = hits_as_struct(got)
    rt, ru, rv, rp = hits_as_struct(ref)
    np.testing.assert_array_equal(gp, rp)
    np.testing.assert_array_equal(gt, rt)
    np.testing.assert_array_equal(gu, ru)
    np.testing.assert_array_equal(gv, rv)
    assert (gp != 0xFFFFFFFF).sum() > 1000
    assert np.all(gp[gp != 0xFFFFFFFF] < 3000)  # duplicates never win a tie


def test_any_matches_brute(oracle):
    tris = random_tris(2000, 3)
    rays = random_rays(3000, 4)
    rays[:, 7] = np.random.default_rng(5).uniform(0.5, 10, 3000)
    b = oracle.bvh(tris)
    np.testing.assert_array_equal(b.trace(rays, any_hit=True), oracle.brute(tris, rays, True))


def test_counts_reported(oracle):
    tris = random_tris(1000, 6)
    rays = random_rays(500, 7)
    b = oracle.bvh(tris)
    b.trace(rays)
    nbox, ntri = b.last_counts
    assert nbox > 500 and ntri > 0
