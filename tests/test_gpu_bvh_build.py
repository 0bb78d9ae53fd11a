"""Device BVH builder: its wide BVH is byte-identical to the host builder's
(same binned SAH, stable partitions, shared collapse) for every width, and
traces the same hits."""
import time

import numpy as np
import pytest

from appearance_amd.raytrace import BUILDER_GPU, BUILDER_HOST, Bvh, rays_bvh_stress
from appearance_amd.scene import generate_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,seed", [("cornell", 0), ("disney", 3), ("bvh_stress", 1)])
def test_device_build_equals_host_build(gpu_available, name, seed):
    s = generate_scene(name, seed=seed)
    for w in (2, 4, 8):
        t0 = time.time()
        h = Bvh(s.positions, width=w, builder=BUILDER_HOST)
        th = time.time() - t0
        t0 = time.time()
        g = Bvh(s.positions, width=w, builder=BUILDER_GPU)
        tg = time.time() - t0
        sh, sg = h.stats(), g.stats()
        print(f"[bvh] {name} W={w}: host {sh['build_ms']:.1f} ms, device {sg['build_ms']:.1f} ms "
              f"(wall {th * 1e3:.0f} / {tg * 1e3:.0f} ms), nodes {sh['n_nodes']} / {sg['n_nodes']}, "
              f"sah {sh['sah_cost']:.4f} / {sg['sah_cost']:.4f}", flush=True)
        assert sh["n_nodes"] == sg["n_nodes"] and sh["depth"] == sg["depth"]
        assert sh["max_stack"] == sg["max_stack"]
        nh, th_ = h.export()
        ng, tg_ = g.export()
        assert np.array_equal(nh, ng)
        assert np.array_equal(th_.view(np.uint32), tg_.view(np.uint32))


def test_device_build_hits(gpu_available):
    s = generate_scene("bvh_stress", seed=1)
    rays = rays_bvh_stress(seed=2, side=256, n_incoherent=1 << 16)
    g = Bvh(s.positions, width=4, builder=BUILDER_GPU)
    h = Bvh(s.positions, width=4, builder=BUILDER_HOST)
    a, b = g.trace_closest(rays), h.trace_closest(rays)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_device_build_rejects_bad_input(gpu_available):
    tris = np.zeros((4, 9), np.float32)
    tris[2, 4] = np.nan
    with pytest.raises(Exception):
        Bvh(tris, builder=BUILDER_GPU)
