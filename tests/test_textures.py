"""Textures, uvs and tangent frames of glTF scenes (CPU): libapt's PNG
decoder against images written by tests/png_util.py (every colour type, bit
depth and filter, Adam7, palettes with transparency) and against an
independent zlib-based decoder on the reference's own images; the material
texture slots of asset.rs:196-345; TEXCOORD_0; tangents given or generated
with the reference's per-vertex method (mesh.rs:82-160) against closed forms."""
from pathlib import Path

import numpy as np
import pytest

from appearance_amd.scene import load_gltf
from tests import png_util
from tests.gltf_util import write_glb

ASSETS = Path("/root/reference/assets")
INVALID = 0xFFFFFFFF


def quad_scene(path, images, materials, uvs=None, tangents=None, textures=None, positions=None, normals=None):
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32) if positions is None else positions
    prim = {"positions": P, "indices": np.array([0, 1, 2, 0, 2, 3], np.uint32), "material": 0,
            "normals": np.tile([0, 0, 1], (4, 1)).astype(np.float32) if normals is None else normals}
    if uvs is not None:
        prim["uvs"] = uvs
    if tangents is not None:
        prim["tangents"] = tangents
    write_glb(path, [{"mesh": 0}], [[prim]], materials, [0], images=images, textures=textures)
    return load_gltf(path)


CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8), (4, 8), (4, 16),
         (6, 8), (6, 16)]


@pytest.mark.parametrize("ctype,depth", CASES)
@pytest.mark.parametrize("interlace", [False, True])
def test_png_decoding(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + (7 if interlace else 0))
    h, w = 13, 21  # not multiples of 8: partial Adam7 passes, padded sub-byte rows
    pal = trns = None
    if ctype == 3:
        n = 1 << depth
        pal = rng.integers(0, 256, (n, 3))
        trns = rng.integers(0, 256, n // 2).tolist()
        img = rng.integers(0, n, (h, w))
    else:
        img = rng.integers(0, 1 << depth, (h, w, png_util.CHANNELS[ctype]))
        if ctype in (0, 2):
            trns = [0] * (2 * png_util.CHANNELS[ctype])
            v = img[3, 4]
            for c in range(png_util.CHANNELS[ctype]):
                trns[2 * c], trns[2 * c + 1] = int(v[c]) >> 8, int(v[c]) & 0xFF
    data = png_util.encode(img, ctype, depth, interlace, palette=pal, trns=trns)
    want = png_util.expected_rgba(img, ctype, depth, pal, trns)
    s = quad_scene(tmp_path / "t.glb", [(data, "image/png")],
                   [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}],
                   uvs=np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32))
    assert len(s.textures) == 1
    np.testing.assert_array_equal(s.textures[0], want)
    if ctype in (0, 2):
        assert (s.textures[0] >> 24)[3, 4] == 0  # the tRNS colour is transparent


def test_png_corrupt_and_other_formats(tmp_path):
    good = png_util.encode(np.zeros((4, 4, 3), np.uint8), 2, 8)
    bad = good[:40] + bytes(len(good) - 40)  # damaged compressed data
    jpeg = b"\xff\xd8\xff\xe0" + bytes(64)
    s = quad_scene(tmp_path / "t.glb", [(bad, "image/png"), (jpeg, "image/jpeg"), (good, "image/png")],
                   [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0},
                                              "metallicRoughnessTexture": {"index": 1}},
                     "emissiveTexture": {"index": 2}}])
    m = s.materials[0]
    assert m.color_texture == INVALID and m.metallic_roughness_texture == INVALID
    assert m.emission_texture == 0 and len(s.textures) == 1


def test_material_texture_slots(tmp_path):
    imgs = [(png_util.encode(np.full((2, 2, 3), k * 20, np.uint8), 2, 8), "image/png") for k in range(10)]
    ext = {"KHR_materials_transmission": {"transmissionFactor": 1.0, "transmissionTexture": {"index": 4}},
           "KHR_materials_clearcoat": {"clearcoatFactor": 1.0, "clearcoatTexture": {"index": 5},
                                       "clearcoatRoughnessTexture": {"index": 6},
                                       "clearcoatNormalTexture": {"index": 7}},
           "KHR_materials_sheen": {"sheenRoughnessTexture": {"index": 8}, "sheenColorTexture": {"index": 9}}}
    mat = {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicRoughnessTexture": {"index": 1}},
           "normalTexture": {"index": 2, "scale": 0.5}, "emissiveTexture": {"index": 3}, "extensions": ext}
    # textures 10 and 11 share image 3: decoded once
    s = quad_scene(tmp_path / "t.glb", imgs, [mat], uvs=np.zeros((4, 2), np.float32),
                   textures=list(range(10)) + [3, 3])
    m = s.materials[0]
    slots = [m.color_texture, m.metallic_roughness_texture, m.normal_texture, m.emission_texture,
             m.transmission_texture, m.clearcoat_texture, m.clearcoat_roughness_texture,
             m.clearcoat_normal_texture, m.sheen_texture, m.sheen_tint_texture]
    assert m.normal_scale == np.float32(0.5)
    assert len(s.textures) == 10 and sorted(slots) == list(range(10))
    for k, slot in enumerate(slots):  # each slot holds its own image
        assert np.all(s.textures[slot] == np.uint32(k * 20 * 0x010101 + 0xFF000000)), k
    # a normal map asks for tangent frames
    assert s.tangents is not None and s.bitangents is not None


def test_uvs_and_generated_tangents(tmp_path):
    """Generated tangents (mesh.rs:82-160) against closed forms: on a planar quad
    with u along +x and v along +y the tangent is +x, the bitangent +y and the
    handedness +1; with u mirrored the tangent is -x and the handedness -1 (the
    bitangent stays +y); a rotated quad rotates the frame."""
    normal_map = (png_util.encode(np.full((2, 2, 3), 128, np.uint8), 2, 8), "image/png")
    mat = [{"normalTexture": {"index": 0}}]
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    s = quad_scene(tmp_path / "a.glb", [normal_map], mat, uvs=uv)
    np.testing.assert_array_equal(s.uvs.reshape(-1, 2), uv[[0, 1, 2, 0, 2, 3]])
    t = s.tangents.reshape(-1, 4)
    np.testing.assert_allclose(t[:, :3], np.tile([1, 0, 0], (6, 1)), atol=1e-6)
    np.testing.assert_array_equal(t[:, 3], 1.0)
    np.testing.assert_allclose(s.bitangents.reshape(-1, 3), np.tile([0, 1, 0], (6, 1)), atol=1e-6)
    uvm = uv.copy()
    uvm[:, 0] = 1 - uvm[:, 0]
    s = quad_scene(tmp_path / "b.glb", [normal_map], mat, uvs=uvm)
    t = s.tangents.reshape(-1, 4)
    np.testing.assert_allclose(t[:, :3], np.tile([-1, 0, 0], (6, 1)), atol=1e-6)
    np.testing.assert_array_equal(t[:, 3], -1.0)
    np.testing.assert_allclose(s.bitangents.reshape(-1, 3), np.tile([0, 1, 0], (6, 1)), atol=1e-6)
    a = np.radians(30)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    P = (np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float64) @ R.T).astype(np.float32)
    N = np.tile(R @ [0, 0, 1], (4, 1)).astype(np.float32)
    s = quad_scene(tmp_path / "c.glb", [normal_map], mat, uvs=uv, positions=P, normals=N)
    np.testing.assert_allclose(s.tangents.reshape(-1, 4)[:, :3], np.tile(R @ [1, 0, 0], (6, 1)), atol=1e-6)
    np.testing.assert_allclose(s.bitangents.reshape(-1, 3), np.tile(R @ [0, 1, 0], (6, 1)), atol=1e-6)


def test_given_tangents_are_kept(tmp_path):
    normal_map = (png_util.encode(np.full((2, 2, 3), 128, np.uint8), 2, 8), "image/png")
    tg = np.array([[0, 1, 0, -1]] * 4, np.float32)
    s = quad_scene(tmp_path / "t.glb", [normal_map], [{"normalTexture": {"index": 0}}],
                   uvs=np.zeros((4, 2), np.float32), tangents=tg)
    t = s.tangents.reshape(-1, 4)
    np.testing.assert_array_equal(t, np.tile(tg[0], (6, 1)))
    # bitangent = cross(n, t) * w = cross(z, y) * -1 = +x
    np.testing.assert_allclose(s.bitangents.reshape(-1, 3), np.tile([1, 0, 0], (6, 1)), atol=1e-7)


def test_untextured_scene_has_no_attribs(tmp_path):
    s = quad_scene(tmp_path / "t.glb", [], [{"pbrMetallicRoughness": {"baseColorFactor": [1, 0, 0, 1]}}])
    assert s.uvs is None and s.tangents is None and not s.textures


@pytest.mark.skipif(not ASSETS.exists(), reason="the reference's assets are not on this host")
@pytest.mark.parametrize("name", ["Duck.glb", "test_models/AttenuationTest.glb"])
def test_reference_images_match_independent_decoder(name):
    """The reference's own PNGs: libapt's texels equal Python's zlib + a numpy
    unfilter, image by image (textures in the order the materials use them)."""
    import json
    import struct
    b = (ASSETS / name).read_bytes()
    jl = struct.unpack("<I", b[12:16])[0]
    doc = json.loads(b[20:20 + jl])
    binc = b[20 + jl + 8:]
    s = load_gltf(ASSETS / name)
    decoded = []
    for im in doc["images"]:
        bv = doc["bufferViews"][im["bufferView"]]
        data = binc[bv.get("byteOffset", 0):bv.get("byteOffset", 0) + bv["byteLength"]]
        decoded.append(png_util.decode(data))
    found = 0
    for t in s.textures:
        assert any(t.shape == d.shape and np.array_equal(t, d) for d in decoded)
        found += 1
    assert found >= 1
