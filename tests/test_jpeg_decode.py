"""Baseline JPEG texture decoding (CPU): libapt's decoder against
libjpeg-turbo's default decompression, through Pillow, which bundles it.

glTF assets can carry JPEG images (ClearCoatTest.glb's clearcoat normal map
is a 1024x1024 4:2:0 JPEG); the reference decodes them with the image crate
(asset.rs:382-420).  Every case must give libjpeg-turbo's samples exactly:
the accurate integer IDCT and its range limit, fancy upsampling (and plain
replication for chroma planes two samples wide or less), fixed-point
YCbCr -> RGB, restart intervals, optimised Huffman tables, grey images."""
from __future__ import annotations

import io
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from appearance_amd._ffi import AptError
from appearance_amd.scene import decode_image, load_gltf

ASSETS = Path("/root/reference/assets")


def _has_pillow_jpeg() -> bool:
    try:
        from PIL import features
        return bool(features.check_codec("jpg"))
    except Exception:
        return False


pytestmark = pytest.mark.skipif(not _has_pillow_jpeg(), reason="Pillow without JPEG support")


def _encode(img, **kw) -> bytes:
    from PIL import Image
    b = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(img)).save(b, "JPEG", **kw)
    return b.getvalue()


def _decode_ref(data: bytes, mode="RGB") -> np.ndarray:
    from PIL import Image
    return np.asarray(Image.open(io.BytesIO(data)).convert(mode))


def _images():
    rng = np.random.default_rng(0)
    y, x = np.mgrid[0:200, 0:300]
    smooth = np.dstack([128 + 100 * np.sin(x / 13.0), 128 + 90 * np.cos(y / 9.0), (x + y) % 256]).astype(np.uint8)
    return {"smooth": smooth, "noise": rng.integers(0, 256, (200, 300, 3), dtype=np.uint8)}


SIZES = [(300, 200), (1, 1), (2, 7), (3, 5), (4, 4), (15, 9), (17, 33), (64, 64)]


@pytest.mark.parametrize("subsampling", [0, 1, 2], ids=["444", "422", "420"])
@pytest.mark.parametrize("quality", [50, 95, 100])
def test_matches_libjpeg_turbo(subsampling, quality):
    for name, img in _images().items():
        for w, h in SIZES:
            d = _encode(img[:h, :w], quality=quality, subsampling=subsampling)
            got = decode_image(d)
            assert got.shape == (h, w, 4)
            assert np.all(got[..., 3] == 255)
            np.testing.assert_array_equal(got[..., :3], _decode_ref(d), err_msg=f"{name} {w}x{h}")


def test_grey_restart_and_optimised_tables():
    img = _images()["noise"][:70, :90]
    grey = _images()["smooth"][:, :, 0]
    for d in (_encode(grey, quality=90), _encode(grey[:3, :7], quality=75)):
        got = decode_image(d)
        np.testing.assert_array_equal(got[..., 0], _decode_ref(d, "L"))
        assert np.array_equal(got[..., 0], got[..., 1]) and np.array_equal(got[..., 0], got[..., 2])
    for kw in (dict(restart_marker_blocks=3), dict(restart_marker_rows=1), dict(optimize=True)):
        d = _encode(img, quality=90, **kw)
        if "optimize" not in kw:
            assert b"\xff\xdd" in d  # a DRI segment
        np.testing.assert_array_equal(decode_image(d)[..., :3], _decode_ref(d))


def test_refused_files():
    img = _images()["noise"][:32, :32]
    with pytest.raises(AptError, match="progressive"):
        decode_image(_encode(img, progressive=True))
    d = _encode(img, quality=90)
    with pytest.raises(AptError):
        decode_image(d[: len(d) // 2] + b"\xff\xd9")  # scan cut short
    with pytest.raises(AptError):
        decode_image(b"\xff\xd8\xff\xe0" + bytes(64))


def test_device_encoder_files_decode_like_libjpeg_turbo(oracle):
    """The oracle's block files (the bytes the device encoder writes) decode
    to libjpeg-turbo's samples too."""
    from tests._oracle import jpeg_block
    from tests.jpeg_util import sample_blocks
    for name, b in sample_blocks().items():
        d = jpeg_block(oracle, b, 95)
        np.testing.assert_array_equal(decode_image(d)[..., :3], _decode_ref(d), err_msg=name)


@pytest.mark.skipif(not ASSETS.exists(), reason="the reference's assets are not on this host")
def test_clearcoat_asset_jpeg_texture():
    """ClearCoatTest.glb: its JPEG clearcoat normal map is decoded (libjpeg-turbo's
    samples) and the CoatNorm material points at it."""
    path = ASSETS / "test_models" / "ClearCoatTest.glb"
    b = path.read_bytes()
    jl = struct.unpack("<I", b[12:16])[0]
    doc = json.loads(b[20:20 + jl])
    binc = b[20 + jl + 8:]
    im = doc["images"][4]
    assert im["mimeType"] == "image/jpeg"
    bv = doc["bufferViews"][im["bufferView"]]
    want = _decode_ref(binc[bv.get("byteOffset", 0):bv.get("byteOffset", 0) + bv["byteLength"]])
    s = load_gltf(path)
    hits = [t for t in s.textures if t.shape == (1024, 1024)
            and np.array_equal(np.stack([(t >> (8 * c)) & 0xFF for c in range(3)], -1).astype(np.uint8), want)]
    assert len(hits) == 1
    k = next(i for i, t in enumerate(s.textures) if t is hits[0])
    assert any(m.clearcoat_normal_texture == k for m in s.materials)
