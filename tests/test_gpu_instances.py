"""Spawned models as a two-level scene in the renderer (SpawnModel /
TransformModel / DestroyModel, scene_resources/mod.rs:115-276): one BLAS per
asset in object space, a TLAS over the instances rebuilt on every world
action.

* Several transformed models (two sharing one asset) render like the FP32
  oracle on the same world-space triangles (numpy placement), with the plain
  and with the ReSTIR integrator.
* A move at a million triangles rebuilds the scene in about a millisecond
  (the TLAS and the world-space attributes only; the BLASes stay), and the
  frame equals a fresh renderer spawned at the final transforms bit for bit.
"""
import uuid

import numpy as np
import pytest

from appearance_amd import world
from appearance_amd.renderer import PathTracerGpu, default_config, plain_config
from appearance_amd.scene import SceneData, generate_scene
from tests._oracle import OracleRestir, oracle_render

pytestmark = pytest.mark.gpu


def relmse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sum((a - b) ** 2) / max(np.sum(b ** 2), 1e-30))


def xform(angle_deg, axis, scale, t):
    """4x4 (row-major numpy; SpawnModel stores it column-major): rotation, non-uniform scale, move."""
    a = np.radians(angle_deg)
    k = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    R = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
    m = np.eye(4)
    m[:3, :3] = R @ np.diag(scale)
    m[:3, 3] = t
    return m.astype(np.float32)


def place_scene(parts):
    """World-space soup of (SceneData, 4x4) parts with the placement kernel's
    float32 order of operations; materials appended per part; the sky of the
    last part that has one (SpawnModel keeps the last spawned model's sky)."""
    pos, nrm, mids, mats = [], [], [], []
    have_n = all(s.normals is not None for s, _ in parts)
    for s, m in parts:
        p = s.positions.reshape(-1, 3).astype(np.float32)
        w = np.empty_like(p)
        for r in range(3):
            w[:, r] = (m[r, 0] * p[:, 0] + m[r, 1] * p[:, 1]) + m[r, 2] * p[:, 2] + m[r, 3]
        pos.append(w.reshape(-1, 9))
        if have_n:
            nm = np.linalg.inv(m[:3, :3].astype(np.float64)).T.astype(np.float32)
            q = s.normals.reshape(-1, 3).astype(np.float32)
            v = np.empty_like(q)
            for r in range(3):
                v[:, r] = (nm[r, 0] * q[:, 0] + nm[r, 1] * q[:, 1]) + nm[r, 2] * q[:, 2]
            ln = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
            nrm.append(np.where(ln[:, None] > 0, v / ln[:, None], v).reshape(-1, 9))
        base = sum(len(x) for x in mats)
        mids.append(np.asarray(s.material_ids, np.uint32) + base)
        mats.append(list(s.materials))
    sky = parts[0][0].sky
    for s, _ in parts:
        if s.sky.sun_enabled > 0 or max(s.sky.zenith) > 0:
            sky = s.sky
    return SceneData(np.concatenate(pos).astype(np.float32),
                     np.concatenate(nrm).astype(np.float32) if have_n else None,
                     np.concatenate(mids), [m for ms in mats for m in ms], parts[0][0].camera, sky)


def _parts():
    cornell = generate_scene("cornell", seed=0)
    disney = generate_scene("disney", seed=3)
    return [
        ("procedural://cornell?seed=0", cornell, xform(20.0, (0, 1, 0), (1.0, 1.0, 1.0), (-1.3, 0.0, 0.0))),
        ("procedural://cornell?seed=0", cornell, xform(-35.0, (0.2, 1, 0), (0.8, 1.2, 0.9), (1.4, 0.1, -0.3))),
        ("procedural://disney?seed=3", disney, xform(10.0, (1, 0, 0), (0.06, 0.06, 0.06), (0.0, -0.2, 1.2))),
    ]


def _camera():
    from tests.scenes_util import camera
    return camera((0.2, 1.1, 5.0), (0.0, 0.7, 0.0), 55.0)


def test_instanced_models_vs_oracle(gpu_available, oracle):
    parts = _parts()
    soup = place_scene([(s, m) for _, s, m in parts])
    cam = _camera()
    pt = PathTracerGpu(plain_config(sample_count=16, max_bounces=5))
    for k, (path, _, m) in enumerate(parts):
        pt.handle_visible_world_action(world.SpawnModel(path, m, uuid.UUID(int=k + 1)))
    pt.set_camera(cam)
    _, got = pt.render_accum((160, 96), with_rgba=False)
    ref, _ = oracle_render(oracle, soup, 160, 96, spp=16, max_bounces=5, rr_start=3, camera=cam, sky=soup.sky)
    r = relmse(got, ref)
    print(f"[instances] 3 models (2 sharing an asset), plain integrator: mean {got.mean():.4f}, "
          f"relMSE vs oracle {r:.3e}", flush=True)
    assert got.mean() > 0.01
    assert r < 1e-3


def test_instanced_models_restir_vs_oracle(gpu_available, oracle):
    parts = _parts()
    soup = place_scene([(s, m) for _, s, m in parts])
    cam = _camera()
    cfg = default_config()
    pt = PathTracerGpu(cfg)
    for k, (path, _, m) in enumerate(parts):
        pt.handle_visible_world_action(world.SpawnModel(path, m, uuid.UUID(int=k + 1)))
    pt.set_camera(cam)
    orr = OracleRestir(oracle, soup, 128, 80)
    for f in range(3):
        _, got = pt.render_accum((128, 80), with_rgba=False)
        ref = orr.frame(cfg, camera=cam, max_bounces=cfg.max_bounces, frame_idx=f)
        r = relmse(got, ref)
        print(f"[instances] ReSTIR frame {f}: relMSE vs oracle {r:.3e}", flush=True)
        assert r < 1e-3


def test_move_at_a_million_triangles(gpu_available):
    """SpawnModel of the 1,048,576-triangle stress model and a Cornell box, then
    TransformModel of each: the rebuild after a move (ms_scene) keeps the
    BLASes and costs about a millisecond; the frame after the moves equals a
    fresh renderer spawned at the final transforms."""
    cam = world.CameraUpdate(world.look_at((0.0, 2.0, 30.0), (0.0, 0.0, 0.0)), 50.0)
    ua, ub = uuid.UUID(int=11), uuid.UUID(int=12)

    def renderer(ta, tb):
        pt = PathTracerGpu(plain_config(sample_count=1, max_bounces=2))
        pt.handle_visible_world_action(world.SpawnModel("procedural://bvh_stress?seed=1", ta, ua))
        pt.handle_visible_world_action(world.SpawnModel("procedural://cornell?seed=0", tb, ub))
        pt.handle_visible_world_action(cam)
        return pt

    ta0 = xform(0.0, (0, 1, 0), (1, 1, 1), (0, 0, 0))
    tb0 = xform(0.0, (0, 1, 0), (2, 2, 2), (0, 0, 12))
    a = renderer(ta0, tb0)
    a.render_accum((128, 64), with_rgba=False)
    first = a.stats()["ms_scene"]
    moves = []
    for k in range(6):
        tb = xform(15.0 * k, (0, 1, 0), (2, 2, 2), (0.5 * k, 0, 12))
        a.handle_visible_world_action(world.TransformModel(ub, tb))
        a.render_accum((128, 64), with_rgba=False)
        moves.append(a.stats()["ms_scene"])
    ta = xform(25.0, (1, 1, 0), (1.1, 0.9, 1.0), (1.0, -0.5, 0.0))
    a.handle_visible_world_action(world.TransformModel(ua, ta))
    a.set_frame_index(0)
    _, moved = a.render_accum((128, 64), with_rgba=False)
    moves.append(a.stats()["ms_scene"])
    _, ref = renderer(ta, tb).render_accum((128, 64), with_rgba=False)
    print(f"[instances] 1,048,588 triangles: first build {first:.1f} ms, moves "
          f"{', '.join(f'{m:.2f}' for m in moves)} ms (median {np.median(moves):.2f})", flush=True)
    np.testing.assert_array_equal(moved, ref)
    assert np.median(moves) < 1.5
