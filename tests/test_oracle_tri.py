"""Pins the oracle's watertight ray-triangle test (Woop, Benthin, Wald 2013).

Checks: a closed-form known answer (t, u, v), strict tmin/tmax bounds,
two-sidedness, degenerate triangles, agreement with an FP64 ray-plane
solution on random inputs away from edges, and watertightness: rays through
shared edges and vertices of a triangulated grid always hit."""
import numpy as np


def ray(o, d, tmin=0.0, tmax=1e30):
    return np.array([*o, tmin, *d, tmax], np.float32)


TRI = np.array([-1, -1, 0, 1, -1, 0, 0, 1, 0], np.float32)


def test_known_answer(oracle):
    ok, tuv = oracle.tri_test(ray((0, 0, -1), (0, 0, 1)), TRI)
    assert ok
    np.testing.assert_array_equal(tuv, np.array([1.0, 0.25, 0.5], np.float32))


def test_bounds_are_strict(oracle):
    assert not oracle.tri_test(ray((0, 0, -1), (0, 0, 1), 0.0, 1.0), TRI)[0]
    assert not oracle.tri_test(ray((0, 0, -1), (0, 0, 1), 1.0, 2.0), TRI)[0]
    assert oracle.tri_test(ray((0, 0, -1), (0, 0, 1), 0.999, 1.001), TRI)[0]


def test_two_sided_and_behind(oracle):
    assert oracle.tri_test(ray((0, 0, 1), (0, 0, -1)), TRI)[0]
    assert not oracle.tri_test(ray((0, 0, 1), (0, 0, 1)), TRI)[0]  # behind the origin


def test_degenerate(oracle):
    deg = np.array([0, 0, 0, 1, 1, 0, 2, 2, 0], np.float32)
    assert not oracle.tri_test(ray((0.5, 0.5, -1), (0, 0, 1)), deg)[0]


def test_against_fp64(oracle):
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(4000):
        v = rng.uniform(-1, 1, 9).astype(np.float32)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        target = v.reshape(3, 3).mean(axis=0) + rng.uniform(-0.6, 0.6, 3).astype(np.float32)
        d = (target - o).astype(np.float32)
        ok, tuv = oracle.tri_test(ray(o, d), v)
        # FP64 reference: solve o + t d = p0 + u e1 + w e2
        P = v.astype(np.float64).reshape(3, 3)
        e1, e2 = P[1] - P[0], P[2] - P[0]
        A = np.stack([-d.astype(np.float64), e1, e2], axis=1)
        if abs(np.linalg.det(A)) < 1e-6:
            continue
        t, u, w = np.linalg.solve(A, o.astype(np.float64) - P[0])
        margin = min(u, w, 1 - u - w)
        if abs(margin) < 1e-4 or abs(t) < 1e-4:
            continue
        expect = margin > 0 and t > 0
        assert ok == expect
        if ok:
            assert abs(tuv[0] - t) <= 1e-5 * max(1.0, abs(t))
            assert abs(tuv[1] - u) < 1e-4 and abs(tuv[2] - w) < 1e-4
        checked += 1
    assert checked > 2000


def test_watertight_grid(oracle):
    # 8x8 grid of quads in the plane z = 0.37x + 0.11y, each split into two
    # triangles.  Rays aimed at every edge midpoint and vertex must hit.
    n = 8
    def z(x, y):
        return np.float32(0.37) * np.float32(x) + np.float32(0.11) * np.float32(y)
    tris = []
    for i in range(n):
        for j in range(n):
            p = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            P = [np.array([x, y, z(x, y)], np.float32) for x, y in p]
            tris.append(np.concatenate([P[0], P[1], P[2]]))
            tris.append(np.concatenate([P[0], P[2], P[3]]))
    tris = np.array(tris, np.float32)
    targets = []
    for i in range(1, n):
        for j in range(1, n):
            targets += [(i, j), (i + 0.5, j), (i, j + 0.5), (i + 0.5, j + 0.5)]
    rng = np.random.default_rng(4)
    for (x, y) in targets:
        tgt = np.array([x, y, z(x, y)], np.float32)
        for _ in range(4):
            o = (tgt + np.array([*rng.uniform(-2, 2, 2), 5.0])).astype(np.float32)
            d = (tgt - o).astype(np.float32)
            hits = sum(oracle.tri_test(ray(o, d), t)[0] for t in tris)
            assert hits >= 1, (x, y, o)
