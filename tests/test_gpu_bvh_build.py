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


def test_device_build_small_scenes(gpu_available):
    """Tiny triangle counts (a single-leaf root, one or two wide levels, a
    level narrower than its wide nodes) and a scene of exact duplicates:
    the device build and collapse give the host's bytes and statistics."""
    rng = np.random.default_rng(31)
    cases = {n: rng.uniform(-1, 1, (n, 9)).astype(np.float32) for n in (1, 2, 3, 5, 17, 64, 257)}
    cases["dup"] = np.repeat(rng.uniform(-1, 1, (1, 9)).astype(np.float32), 40, axis=0)
    for name, tris in cases.items():
        for w in (2, 4, 8):
            print(f"[bvh] small scene {name} W={w}", flush=True)
            h = Bvh(tris, width=w, builder=BUILDER_HOST)
            g = Bvh(tris, width=w, builder=BUILDER_GPU)
            sh, sg = h.stats(), g.stats()
            for k in ("n_nodes", "n_leaves", "depth", "max_stack"):
                assert sh[k] == sg[k], (name, w, k, sh[k], sg[k])
            nh, th = h.export()
            ng, tg = g.export()
            assert np.array_equal(nh, ng) and np.array_equal(th.view(np.uint32), tg.view(np.uint32))


def test_device_build_rejects_bad_input(gpu_available):
    tris = np.zeros((4, 9), np.float32)
    tris[2, 4] = np.nan
    with pytest.raises(Exception):
        Bvh(tris, builder=BUILDER_GPU)


def decode_nodes(raw: np.ndarray, width: int):
    """Exported wide nodes -> (child refs [n, W], decoded child boxes lo/hi [n, W, 3]) in float64."""
    from appearance_amd.raytrace import NODE_BYTES
    nb = NODE_BYTES[width]
    n = raw.size // nb
    rec = raw.reshape(n, nb)
    org = rec[:, 0:12].copy().view(np.float32).astype(np.float64)          # [n, 3]
    exp = rec[:, 12:15].astype(np.int64) - 127                              # [n, 3]
    child = rec[:, 16:16 + 4 * width].copy().view(np.uint32)                # [n, W]
    q = rec[:, 16 + 4 * width:16 + 10 * width].reshape(n, 6, width).astype(np.float64)
    step = np.ldexp(1.0, exp)                                               # [n, 3]
    lo = org[:, None, :] + np.transpose(q[:, 0:3, :], (0, 2, 1)) * step[:, None, :]
    hi = org[:, None, :] + np.transpose(q[:, 3:6, :], (0, 2, 1)) * step[:, None, :]
    return child, lo, hi


@pytest.mark.parametrize("width", [2, 4, 8])
def test_quantized_boxes_contain_their_subtrees(gpu_available, width):
    """Every decoded child box (origin + q * 2^e, exactly) holds all triangles
    below that child: the quantization rounds outwards."""
    from tests.test_oracle_bvh import random_tris
    tris = np.concatenate([random_tris(3000, 21), generate_scene("cornell").positions])
    b = Bvh(tris, width=width)
    raw, trec = b.export()
    child, lo, hi = decode_nodes(raw, width)
    v = trec.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)   # leaf-order triangles
    tlo, thi = v.min(axis=1), v.max(axis=1)
    n = child.shape[0]
    sub_lo = np.full((n, 3), np.inf)
    sub_hi = np.full((n, 3), -np.inf)
    checked = 0
    for i in range(n - 1, -1, -1):   # breadth-first order: children after parents
        for k in range(width):
            r = int(child[i, k])
            if r == 0xFFFFFFFF:
                continue
            if r & 0x80000000:
                first, cnt = r & 0x0FFFFFFF, ((r >> 28) & 7) + 1
                clo, chi = tlo[first:first + cnt].min(axis=0), thi[first:first + cnt].max(axis=0)
            else:
                assert r > i
                clo, chi = sub_lo[r], sub_hi[r]
            assert np.all(lo[i, k] <= clo) and np.all(hi[i, k] >= chi), (i, k)
            sub_lo[i] = np.minimum(sub_lo[i], clo)
            sub_hi[i] = np.maximum(sub_hi[i], chi)
            checked += 1
    assert checked == n + b.stats()["n_leaves"] - 1
