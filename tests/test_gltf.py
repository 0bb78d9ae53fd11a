"""glTF ingestion (apt_scene_load_gltf, gltf.cpp) against an independent
numpy reader (tests/gltf_util.py): the flattened triangles, vertex normals
(given or generated) and material ids of a written test scene and of the
reference's own .glb assets, the material factor mapping of the reference's
loader (crates/appearance-model/src/asset.rs:199-339), and the error cases."""
from pathlib import Path

import numpy as np
import pytest

from appearance_amd._ffi import AptError
from appearance_amd.scene import load_gltf
from tests.gltf_util import demo_scene, read_glb, write_glb

ASSETS = Path("/root/reference/assets")
APT_ERR_UNSUPPORTED, APT_ERR_NOT_FOUND = 4, 6


def compare(path):
    s = load_gltf(path)
    pos, nrm, mids, doc = read_glb(path)
    assert s.positions.shape == pos.shape
    scale = max(1.0, float(np.abs(pos).max()))
    np.testing.assert_allclose(s.positions, pos, rtol=0, atol=2e-6 * scale)
    np.testing.assert_allclose(s.normals, nrm, rtol=0, atol=2e-5)
    np.testing.assert_array_equal(s.material_ids, mids)
    assert len(s.materials) == max(1, len(doc.get("materials", [])))
    return s, doc


def test_demo_scene_matches_numpy_reader(tmp_path):
    p = tmp_path / "demo.glb"
    demo_scene(p)
    s, _ = compare(p)
    assert s.n_tris == 2 + 12 + 12 + 2 + 2
    # generated normals of the boxes are unit length
    np.testing.assert_allclose(np.linalg.norm(s.normals.reshape(-1, 3), axis=1), 1.0, atol=1e-6)


def test_material_mapping(tmp_path):
    p = tmp_path / "demo.glb"
    demo_scene(p)
    m = load_gltf(p).materials
    f = np.float32
    assert list(m[0].color) == [f(0.7), f(0.7), f(0.65)] and m[0].roughness == f(0.8) and m[0].metallic == 0
    assert m[1].clearcoat == f(0.7) and m[1].clearcoat_roughness == f(0.1)
    assert m[1].sheen == f(0.6) and list(m[1].sheen_tint) == [f(0.2), f(0.3), f(0.4)]
    assert m[1].specular == f(0.5) and list(m[1].specular_tint) == [f(1.0), f(0.9), f(0.8)]
    np.testing.assert_allclose(list(m[2].emission), [12.0, 9.6, 7.2], rtol=1e-6)
    assert m[3].transmission == 1.0 and m[3].eta == f(1.0 / 1.45)
    np.testing.assert_allclose(list(m[3].absorption), [0.05, 0.1, 0.15], rtol=1e-5)
    for mm in m:
        assert mm.alpha_cutoff == f(0.5) and mm.bsdf_model == 0
        assert mm.color_texture == 0xFFFFFFFF and mm.normal_texture == 0xFFFFFFFF
    # materials without the extensions keep the reference's defaults
    assert m[0].eta == f(1.0 / 1.5) and m[0].transmission == 0 and list(m[0].specular_tint) == [1, 1, 1]


@pytest.mark.skipif(not ASSETS.exists(), reason="the reference's assets are not on this host")
@pytest.mark.parametrize("name", ["CornellBoxConor.glb", "EmissivePlane.glb", "GlassPanel.glb", "Duck.glb",
                                  "SponzaOrb.glb", "SponzaNeon2.glb", "test_models/ClearCoatTest.glb",
                                  "test_models/AttenuationTest.glb"])
def test_reference_assets(name):
    s, doc = compare(ASSETS / name)
    assert s.n_tris > 0
    # the glTF's own counts: every triangle primitive of every mesh instance
    n = 0
    for node in doc["nodes"]:
        if "mesh" in node:
            for p in doc["meshes"][node["mesh"]]["primitives"]:
                n += doc["accessors"][p["indices"]]["count"] // 3
    reachable = s.n_tris
    assert reachable <= n


def test_errors(tmp_path):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    p = tmp_path / "lines.glb"
    write_glb(p, [{"mesh": 0}], [[{"positions": v, "indices": [0, 1, 1, 2], "mode": 1}]], [], [0])
    with pytest.raises(AptError) as e:
        load_gltf(p)
    assert e.value.status == APT_ERR_UNSUPPORTED
    with pytest.raises(AptError) as e:
        load_gltf(tmp_path / "missing.glb")
    assert e.value.status == APT_ERR_NOT_FOUND
    q = tmp_path / "truncated.glb"
    write_glb(q, [{"mesh": 0}], [[{"positions": v, "indices": [0, 1, 2]}]], [], [0])
    q.write_bytes(q.read_bytes()[:40])
    with pytest.raises(AptError) as e:
        load_gltf(q)
    assert e.value.status == APT_ERR_UNSUPPORTED
    r = tmp_path / "ok.glb"
    write_glb(r, [{"mesh": 0}], [[{"positions": v, "indices": [0, 1, 2]}]], [], [0])
    s = load_gltf(r)
    assert s.n_tris == 1 and len(s.materials) == 1 and s.materials[0].metallic == 1.0


CONOR = Path(__file__).resolve().parent / "golden" / "cornell_conor.glb"


def test_cornell_conor_fixture_loads():
    """The committed re-encoding of the reference's CornellBoxConor.glb
    (tests/golden/make_cornell_conor.py): 36 triangles in a 2 m box, eight
    materials, the light at emissiveFactor 1 x emissive strength 5."""
    s, doc = compare(CONOR)
    assert s.n_tris == 36 and len(s.materials) == 8
    lo, hi = s.positions.reshape(-1, 3).min(0), s.positions.reshape(-1, 3).max(0)
    np.testing.assert_allclose(lo, [-1, 0, -1], atol=1e-5)
    np.testing.assert_allclose(hi, [1, 2, 1], atol=1e-5)
    light = [i for i, m in enumerate(s.materials) if max(m.emission) > 0]
    assert light == [5] and list(s.materials[5].emission) == [5.0, 5.0, 5.0]
    assert (s.material_ids == 5).sum() == 2


@pytest.mark.skipif(not (ASSETS / "CornellBoxConor.glb").exists(), reason="the reference's assets are not on this host")
def test_cornell_conor_fixture_equals_original():
    """libapt reads the original asset and the fixture into the same triangles
    and materials, up to the order of the triangles and the one texture the
    fixture drops (the light's 1024x1024 emission image)."""
    a = load_gltf(ASSETS / "CornellBoxConor.glb")
    b = load_gltf(CONOR)
    order = np.argsort(a.material_ids, kind="stable")
    np.testing.assert_allclose(b.positions, a.positions[order], atol=1e-6)
    np.testing.assert_allclose(b.normals, a.normals[order], atol=1e-6)
    np.testing.assert_array_equal(b.material_ids, a.material_ids[order])
    assert [t.shape for t in a.textures] == [(1024, 1024)] and not b.textures
    assert a.materials[5].emission_texture == 0
    for m in a.materials:
        for f, _ in m._fields_:
            if f.endswith("_texture"):
                setattr(m, f, 0xFFFFFFFF)
    assert [bytes(m) for m in a.materials] == [bytes(m) for m in b.materials]
