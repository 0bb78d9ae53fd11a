"""Textured scenes on the GPU against the FP32 oracle (material textures,
uvs and tangent frames of material_pool_bindings.wgsl:15-154 and
vertex_pool_bindings.wgsl:43-58): a room of glTF materials that uses every
texture slot, read by libapt's glTF reader (PNG images, generated tangents),

* set as an explicit scene, with the plain and with the ReSTIR integrator;
* spawned twice as a model (SpawnModel of the GLB, one copy rotated and
  scaled non-uniformly): the two-level scene with per-asset textures, against
  the oracle on the same world-space triangles, uvs and tangent frames placed
  in numpy, with the plain integrator and (under a sky panorama) with the
  ReSTIR one.
"""
import uuid

import numpy as np
import pytest

from appearance_amd import world
from appearance_amd._ffi import Material
from appearance_amd.renderer import PathTracerGpu, default_config, plain_config
from appearance_amd.scene import SceneData, load_gltf
from tests._oracle import OracleRestir, _unit_rows, oracle_render
from tests.gltf_util import textured_room
from tests.scenes_util import camera

pytestmark = pytest.mark.gpu


def relmse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sum((a - b) ** 2) / max(np.sum(b ** 2), 1e-30))


def _cam():
    return camera((0.3, 1.4, 3.6), (0.0, 1.0, -2.0), 60.0)


@pytest.fixture(scope="module")
def room(tmp_path_factory):
    path = tmp_path_factory.mktemp("tex") / "room.glb"
    textured_room(path)
    return path, load_gltf(path)


def test_textured_room_plain_vs_oracle(gpu_available, oracle, room):
    _, s = room
    assert len(s.textures) == 10 and s.tangents is not None
    pt = PathTracerGpu(plain_config(sample_count=16, max_bounces=5))
    pt.set_scene(s)
    pt.set_camera(_cam())
    _, got = pt.render_accum((160, 120), with_rgba=False)
    ref, _ = oracle_render(oracle, s, 160, 120, spp=16, max_bounces=5, rr_start=3, camera=_cam())
    r = relmse(got, ref)
    print(f"[textures] room, plain integrator: mean {got.mean():.4f} (oracle {ref.mean():.4f}), "
          f"relMSE {r:.3e}", flush=True)
    assert np.isfinite(got).all() and got.mean() > 0.05
    assert r < 1e-4


def test_textured_room_jpeg_images_plain_vs_oracle(gpu_available, oracle, tmp_path):
    """The same room with its RGB images stored as JPEG (decoded by libapt's
    baseline decoder as the glTF is read) renders like the oracle on the
    decoded texels."""
    path = tmp_path / "room_jpeg.glb"
    textured_room(path, jpeg=True)
    s = load_gltf(path)
    assert len(s.textures) == 10
    pt = PathTracerGpu(plain_config(sample_count=16, max_bounces=5))
    pt.set_scene(s)
    pt.set_camera(_cam())
    _, got = pt.render_accum((160, 120), with_rgba=False)
    ref, _ = oracle_render(oracle, s, 160, 120, spp=16, max_bounces=5, rr_start=3, camera=_cam())
    r = relmse(got, ref)
    print(f"[textures] room with JPEG images, plain integrator: relMSE {r:.3e}", flush=True)
    assert np.isfinite(got).all() and got.mean() > 0.05
    assert r < 1e-4


def test_textured_room_restir_vs_oracle(gpu_available, oracle, room):
    _, s = room
    cfg = default_config()
    pt = PathTracerGpu(cfg)
    pt.set_scene(s)
    pt.set_camera(_cam())
    orr = OracleRestir(oracle, s, 128, 96)
    for f in range(3):
        _, got = pt.render_accum((128, 96), with_rgba=False)
        ref = orr.frame(cfg, camera=_cam(), max_bounces=cfg.max_bounces, frame_idx=f)
        r = relmse(got, ref)
        print(f"[textures] room, ReSTIR frame {f}: relMSE {r:.3e}", flush=True)
        assert np.isfinite(got).all()
        assert r < 1e-3


def _xform(angle_deg, axis, scale, t):
    a = np.radians(angle_deg)
    k = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    R = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
    m = np.eye(4)
    m[:3, :3] = R @ np.diag(scale)
    m[:3, 3] = t
    return m.astype(np.float32)


def _place(s, m, mat_base):
    """One spawned copy in world space with the placement kernel's float32
    operations: positions by M, normals, tangents and bitangents by the
    inverse transpose of its 3x3 part and renormalised per vertex."""
    nm = np.linalg.inv(m[:3, :3].astype(np.float64)).T.astype(np.float32)

    def lin(mm, v, add=None):
        out = np.empty_like(v)
        for r in range(3):
            out[:, r] = (mm[r, 0] * v[:, 0] + mm[r, 1] * v[:, 1]) + mm[r, 2] * v[:, 2]
            if add is not None:
                out[:, r] += add[r]
        return out

    p = s.positions.reshape(-1, 3).astype(np.float32)
    pos = lin(m, p, m[:3, 3]).reshape(-1, 9)
    nrm = _unit_rows(lin(nm, s.normals.reshape(-1, 3).astype(np.float32))).reshape(-1, 9)
    q = s.tangents.reshape(-1, 4).astype(np.float32)
    b = s.bitangents.reshape(-1, 3).astype(np.float32)
    tan = _unit_rows(lin(nm, q[:, :3])).reshape(-1, 9)
    bit = _unit_rows(lin(nm, b)).reshape(-1, 9)
    return pos, nrm, tan, bit, s.uvs, np.asarray(s.material_ids, np.uint32) + mat_base


def test_spawned_textured_models_vs_oracle(gpu_available, oracle, room):
    path, s = room
    ms = [_xform(0.0, (0, 1, 0), (1, 1, 1), (0, 0, 0)),
          _xform(30.0, (0.3, 1, 0.1), (0.25, 0.35, 0.3), (0.4, 0.8, 1.2))]
    parts = [_place(s, m, k * len(s.materials)) for k, m in enumerate(ms)]
    soup = SceneData(np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]),
                     np.concatenate([p[5] for p in parts]),
                     [Material.from_buffer_copy(m) for m in list(s.materials) * 2], s.camera, s.sky,
                     uvs=np.concatenate([p[4] for p in parts]),
                     tangents=np.zeros((sum(len(p[0]) for p in parts), 12), np.float32),
                     textures=list(s.textures) * 2)
    # the second copy's materials point at its own textures (each asset carries its images)
    for m in soup.materials[len(s.materials):]:
        for f, _ in m._fields_:
            if f.endswith("_texture") and getattr(m, f) != 0xFFFFFFFF:
                setattr(m, f, getattr(m, f) + len(s.textures))
    frames = (np.concatenate([p[2] for p in parts]), np.concatenate([p[3] for p in parts]))
    pt = PathTracerGpu(plain_config(sample_count=16, max_bounces=5))
    for k, m in enumerate(ms):
        pt.handle_visible_world_action(world.SpawnModel(str(path), m, uuid.UUID(int=k + 1)))
    pt.set_camera(_cam())
    _, got = pt.render_accum((160, 120), with_rgba=False)
    ref, _ = oracle_render(oracle, soup, 160, 120, spp=16, max_bounces=5, rr_start=3, camera=_cam(),
                           sky=s.sky, frames=frames)
    r = relmse(got, ref)
    print(f"[textures] two spawned rooms (one rotated, scaled 0.25/0.35/0.3): mean {got.mean():.4f}, "
          f"relMSE {r:.3e}", flush=True)
    assert np.isfinite(got).all() and got.mean() > 0.05
    assert r < 1e-4


def test_spawned_textured_models_restir_vs_oracle(gpu_available, oracle, room):
    """The ReSTIR integrator on the two spawned rooms (the textured two-level
    kernel variants) under a sky panorama."""
    path, s = room
    ms = [_xform(0.0, (0, 1, 0), (1, 1, 1), (0, 0, 0)),
          _xform(30.0, (0.3, 1, 0.1), (0.25, 0.35, 0.3), (0.4, 0.8, 1.2))]
    parts = [_place(s, m, k * len(s.materials)) for k, m in enumerate(ms)]
    soup = SceneData(np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]),
                     np.concatenate([p[5] for p in parts]),
                     [Material.from_buffer_copy(m) for m in list(s.materials) * 2], s.camera, s.sky,
                     uvs=np.concatenate([p[4] for p in parts]),
                     tangents=np.zeros((sum(len(p[0]) for p in parts), 12), np.float32),
                     textures=list(s.textures) * 2)
    for m in soup.materials[len(s.materials):]:
        for f, _ in m._fields_:
            if f.endswith("_texture") and getattr(m, f) != 0xFFFFFFFF:
                setattr(m, f, getattr(m, f) + len(s.textures))
    frames = (np.concatenate([p[2] for p in parts]), np.concatenate([p[3] for p in parts]))
    j, i = np.mgrid[0:16, 0:32].astype(np.float32)
    pano = np.stack([0.2 + 0.02 * i, 0.3 + 0.03 * j, 0.6 + 0.0 * i], -1).astype(np.float32)
    cfg = default_config()
    pt = PathTracerGpu(cfg)
    for k, m in enumerate(ms):
        pt.handle_visible_world_action(world.SpawnModel(str(path), m, uuid.UUID(int=k + 1)))
    pt.set_camera(_cam())
    pt.set_sky_panorama(pano)
    orr = OracleRestir(oracle, soup, 128, 96, frames=frames, sky_pano=pano)
    for f in range(3):
        _, got = pt.render_accum((128, 96), with_rgba=False)
        ref = orr.frame(cfg, camera=_cam(), sky=s.sky, max_bounces=cfg.max_bounces, frame_idx=f)
        r = relmse(got, ref)
        print(f"[textures] two spawned rooms under a panorama, ReSTIR frame {f}: relMSE {r:.3e}", flush=True)
        assert np.isfinite(got).all()
        assert r < 1e-3
