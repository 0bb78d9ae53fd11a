"""Traversal stack guard on the GPU: a tree deeper than the kernels' stack
(tests.scenes_util.deep_tree_triangles at widths 4 and 8) is refused with
APT_ERR_UNSUPPORTED by the ray queries, the renderer, the server and the
autotuner instead of being traced with subtrees lost; at width 2 the same
triangles fit and trace bit-exactly."""
import numpy as np
import pytest

from appearance_amd._ffi import AptError
from appearance_amd.raytrace import RAY_DTYPE, Bvh
from appearance_amd.renderer import PathTracerGpu, plain_config
from appearance_amd.serving import Server
from tests._oracle import hits_as_struct
from tests.scenes_util import deep_tree_scene, deep_tree_triangles

pytestmark = pytest.mark.gpu
APT_ERR_UNSUPPORTED = 4


def _rays(tris, n, seed):
    """Rays along -z aimed at points of the first 60 triangles (a quarter
    offset below them, so they miss)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 60, n)
    a, b = rng.random(n), rng.random(n)
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    t = tris[k].astype(np.float64)
    p = t[:, 0:3] + a[:, None] * (t[:, 3:6] - t[:, 0:3]) + b[:, None] * (t[:, 6:9] - t[:, 0:3])
    p[rng.random(n) < 0.25, 1] -= 0.5
    r = np.zeros(n, RAY_DTYPE)
    r["org"][:, 0] = p[:, 0]
    r["org"][:, 1] = p[:, 1]
    r["org"][:, 2] = 1.0
    r["dir"][:, 2] = -1.0
    r["tmax"] = 1e30
    return r


def test_ray_queries(gpu_available, oracle):
    tris = deep_tree_triangles()
    rays = _rays(tris, 4096, 3)
    for w in (4, 8):
        b = Bvh(tris, width=w)
        assert b.stats()["max_stack"] > 72
        with pytest.raises(AptError) as e:
            b.trace_closest(rays)
        assert e.value.status == APT_ERR_UNSUPPORTED
        with pytest.raises(AptError):
            b.trace_any(rays)
    b = Bvh(tris, width=2)
    got = b.trace_closest(rays)
    t, _, _, prim = hits_as_struct(oracle.brute(tris, rays.view(np.float32).reshape(-1, 8)))
    np.testing.assert_array_equal(got["prim"], prim)
    assert (prim != 0xFFFFFFFF).sum() > 2000


def test_renderer_refuses_then_renders_at_width_2(gpu_available):
    s = deep_tree_scene()
    pt = PathTracerGpu(plain_config(sample_count=1, max_bounces=2))
    pt.set_scene(s)  # the scene is built at the next render
    pt.set_camera(s.camera)
    with pytest.raises(AptError) as e:
        pt.render_accum((64, 64), with_rgba=False)
    assert e.value.status == APT_ERR_UNSUPPORTED
    pt2 = PathTracerGpu(plain_config(sample_count=1, max_bounces=2, bvh_width=2))
    pt2.set_scene(s)
    pt2.set_camera(s.camera)
    _, img = pt2.render_accum((64, 64), with_rgba=False)
    assert np.isfinite(img).all()


def test_server_refuses(gpu_available):
    srv = Server(path_slots=1 << 16, bvh_width=4)
    with pytest.raises(AptError) as e:
        srv.add_scene(deep_tree_scene())
    assert e.value.status == APT_ERR_UNSUPPORTED
    srv.close()


def test_autotuner_skips_widths_that_do_not_fit(gpu_available):
    from appearance_amd.autotune import tune
    rep = tune(deep_tree_scene(), probe=(64, 64), spp=1, bounces=2, repeats=1)
    rejected = {c["bvh_width"] for c in rep["candidates"] if "rejected" in c}
    assert rejected == {4, 8}
    assert rep["best"]["bvh_width"] == 2
