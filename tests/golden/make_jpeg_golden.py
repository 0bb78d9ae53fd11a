"""Writes tests/golden/jpeg_pillow_q95.npz: Pillow's JPEG files (its bundled
libjpeg-turbo, 4:2:0, quality 95 as the reference node's
turbojpeg::compress(image, 95, Sub2x2), node.rs:65-79) for the sample blocks
of tests/jpeg_util.py, so the oracle stays pinned to the library's bytes
even where another Pillow build is installed.

Usage: python tests/golden/make_jpeg_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from tests.jpeg_util import pillow_jpeg, sample_blocks  # noqa: E402


def main() -> None:
    from PIL import features
    blocks = sample_blocks()
    names = sorted(blocks)
    files = [pillow_jpeg(blocks[n], 95) for n in names]
    offs = np.zeros(len(files) + 1, np.int64)
    offs[1:] = np.cumsum([len(f) for f in files])
    np.savez_compressed(ROOT / "tests" / "golden" / "jpeg_pillow_q95.npz",
                        names=np.array(names), data=np.frombuffer(b"".join(files), np.uint8), offsets=offs,
                        libjpeg_turbo=np.array(str(features.version_feature("libjpeg_turbo"))))
    print(f"{len(files)} files, {offs[-1]} bytes")


if __name__ == "__main__":
    main()
