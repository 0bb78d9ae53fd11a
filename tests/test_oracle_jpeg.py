"""The oracle's JPEG block compression is pinned to libjpeg-turbo's bytes.

The reference node compresses each 64x64 RGBX block with
turbojpeg::compress(image, 95, Sub2x2) (node.rs:65-79); the turbojpeg crate
wraps libjpeg-turbo.  Pillow bundles the same library, so its 4:2:0 output at
each quality is the ground truth: the oracle must reproduce it byte for byte
(colour conversion, downsampling bias, integer DCT, rounding, Huffman coding,
stuffing and markers all show up in the bytes).  A committed fixture keeps
the q95 files of libjpeg-turbo 3.1.4 for installs with another build."""
from __future__ import annotations

import io
from pathlib import Path

import numpy as np
import pytest

from tests._oracle import jpeg_block
from tests.jpeg_util import pillow_jpeg, random_blocks, sample_blocks

GOLDEN = Path(__file__).resolve().parent / "golden" / "jpeg_pillow_q95.npz"
QUALITIES = [1, 2, 10, 25, 49, 50, 51, 75, 90, 95, 99, 100]


def _has_pillow_jpeg() -> bool:
    try:
        from PIL import features
        return bool(features.check_codec("jpg"))
    except Exception:
        return False


needs_pillow = pytest.mark.skipif(not _has_pillow_jpeg(), reason="Pillow without JPEG support")


def _scan(f: bytes) -> bytes:
    """The entropy-coded data between SOS and EOI."""
    sos = f.index(b"\xff\xda")
    n = int.from_bytes(f[sos + 2:sos + 4], "big")
    return f[sos + 2 + n:-2]


def test_fixture_matches_oracle(oracle):
    z = np.load(GOLDEN, allow_pickle=False)
    blocks = sample_blocks()
    offs = z["offsets"]
    names = [str(n) for n in z["names"]]
    assert sorted(blocks) == names
    for i, name in enumerate(names):
        want = z["data"][offs[i]:offs[i + 1]].tobytes()
        assert jpeg_block(oracle, blocks[name], 95) == want, name


@needs_pillow
@pytest.mark.parametrize("quality", QUALITIES)
def test_sample_blocks_match_libjpeg_turbo(oracle, quality):
    for name, b in sample_blocks().items():
        assert jpeg_block(oracle, b, quality) == pillow_jpeg(b, quality), (name, quality)


@needs_pillow
@pytest.mark.parametrize("quality", [95, 80])
def test_random_blocks_match_libjpeg_turbo(oracle, quality):
    for i, b in enumerate(random_blocks(48, seed=quality)):
        assert jpeg_block(oracle, b, quality) == pillow_jpeg(b, quality), (i, quality)


def test_structure_stuffing_and_x_channel(oracle):
    blocks = sample_blocks()
    files = {n: jpeg_block(oracle, b, 95) for n, b in blocks.items()}
    for n, f in files.items():
        assert f[:2] == b"\xff\xd8" and f[-2:] == b"\xff\xd9", n
        s = _scan(f)
        # every 0xFF in the scan is followed by a stuffed zero
        i = s.find(b"\xff")
        while i >= 0:
            assert s[i + 1] == 0, n
            i = s.find(b"\xff", i + 2)
    assert any(b"\xff\x00" in _scan(f) for f in files.values())
    # the X byte of RGBX is ignored
    b = blocks["smooth_noise"].copy()
    b[..., 3] = 0
    assert jpeg_block(oracle, b, 95) == files["smooth_noise"]


@needs_pillow
def test_decodes_close_to_input(oracle):
    from PIL import Image
    for name in ("ramp", "smooth_noise", "flat"):
        b = sample_blocks()[name]
        img = np.asarray(Image.open(io.BytesIO(jpeg_block(oracle, b, 95))).convert("RGB"), np.float64)
        mse = np.mean((img - b[..., :3].astype(np.float64)) ** 2)
        psnr = 10 * np.log10(255.0 ** 2 / max(mse, 1e-12))
        assert img.shape == (64, 64, 3)
        assert psnr > 30.0, (name, psnr)  # 4:2:0 chroma on noisy content: ~33 dB
