"""Traversal stack guard (CPU side).

The traversal kernels hold apt_trace_stack_capacity() stack entries (8 in
LDS at the least plus 64 private ones).  A tree whose worst-case stack is
deeper would lose subtrees when the stack fills, so every entry point that
traces refuses it with APT_ERR_UNSUPPORTED; apt_bvh_plan reports the same
verdict on the host, without a device.  The GPU side (renderer, server,
autotuner, ray queries) is in test_gpu_stack_guard.py."""
import numpy as np

from appearance_amd.raytrace import plan, stack_capacity
from tests.scenes_util import deep_tree_triangles

APT_ERR_UNSUPPORTED = 4


def test_capacity():
    assert stack_capacity() == 8 + 64


def test_deep_tree_is_refused_at_widths_4_and_8():
    tris = deep_tree_triangles()
    for w in (4, 8):
        rc, st = plan(tris, width=w)
        assert rc == APT_ERR_UNSUPPORTED, (w, st)
        assert st["max_stack"] > stack_capacity()
        assert st["n_tris"] == len(tris)
    rc, st = plan(tris, width=2)
    assert rc == 0 and st["max_stack"] <= stack_capacity()


def test_ordinary_scene_fits():
    rng = np.random.default_rng(5)
    c = rng.uniform(-10, 10, (20000, 1, 3))
    tris = (c + rng.normal(0, 0.1, (20000, 3, 3))).reshape(-1, 9).astype(np.float32)
    for w in (2, 4, 8):
        rc, st = plan(tris, width=w)
        assert rc == 0 and 1 <= st["max_stack"] <= stack_capacity(), (w, st)
