"""JPEG compression of 64x64 render-node blocks on the device.

The reference node compresses each 64x64 RGBX block it sends with
turbojpeg::compress(image, 95, Subsamp::Sub2x2)
(crates/appearance-render-loop/src/node.rs:65-79, NODE_PIXEL_FORMAT RGBX,
host.rs:22).  compress() and compress_blocks() are that call for one block or
a batch, run by libapt's kernels; every file is byte for byte what
libjpeg-turbo's default compression writes.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import check, lib

QUALITY = 95          # node.rs:76
FILE_MAX = 41600      # APT_JPEG_FILE_MAX: bytes one block's file can take
MAX_BLOCKS = 16384    # APT_JPEG_MAX_BLOCKS


def compress_blocks(blocks, quality: int = QUALITY, timing: bool = False):
    """Compress (n, 64, 64, 4) uint8 RGBX blocks (or their bytes, n * 16384)
    into n JFIF files.  With timing=True returns (files, device ms)."""
    a = np.ascontiguousarray(np.frombuffer(bytes(blocks), np.uint8) if isinstance(blocks, (bytes, bytearray))
                             else blocks, np.uint8).reshape(-1)
    if a.size == 0 or a.size % (64 * 64 * 4):
        raise ValueError(f"compress_blocks: {a.size} bytes is not a whole number of 64x64 RGBX blocks")
    n = a.size // (64 * 64 * 4)
    out = np.empty(n * FILE_MAX, np.uint8)
    offs = (C.c_uint64 * (n + 1))()
    ms = C.c_float(0.0)
    check(lib().apt_jpeg_compress_blocks(C.c_void_p(a.ctypes.data), n, int(quality), C.c_void_p(out.ctypes.data),
                                         out.size, offs, C.byref(ms)), "apt_jpeg_compress_blocks")
    files = [out[offs[b]:offs[b + 1]].tobytes() for b in range(n)]
    return (files, ms.value) if timing else files


def compress(image, quality: int = QUALITY) -> bytes:
    """turbojpeg::compress(image, quality, Sub2x2) of one 64x64 RGBX block."""
    a = np.asarray(image, np.uint8)
    if a.shape != (64, 64, 4):
        raise ValueError(f"compress: expected a (64, 64, 4) RGBX block, got {a.shape}")
    return compress_blocks(a[None], quality)[0]
