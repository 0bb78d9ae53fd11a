"""Radiance HDR (RGBE) files for the tests, independent of libapt's decoder:
Greg Ward's shared-exponent encoding (frexp of the largest channel), written
with flat, run-length (new style) or repeat-marker (old style) scanlines, and
the texel values those bytes stand for, c * 2^(e - 136)."""
from __future__ import annotations

import math

import numpy as np


def to_rgbe(rgb: np.ndarray) -> np.ndarray:
    """(h, w, 3) float -> (h, w, 4) uint8 RGBE."""
    rgb = np.asarray(rgb, np.float64)
    out = np.zeros(rgb.shape[:2] + (4,), np.uint8)
    for idx in np.ndindex(rgb.shape[:2]):
        v = rgb[idx].max()
        if v < 1e-32:
            continue
        m, e = math.frexp(v)
        scale = m * 256.0 / v
        out[idx] = [min(255, int(rgb[idx][0] * scale)), min(255, int(rgb[idx][1] * scale)),
                    min(255, int(rgb[idx][2] * scale)), e + 128]
    return out


def from_rgbe(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, np.uint8)
    e = q[..., 3].astype(np.int64)
    f = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return (q[..., :3].astype(np.float64) * f[..., None]).astype(np.float32)


def _rle_channel(vals: np.ndarray) -> bytes:
    out = bytearray()
    i, n = 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), int(vals[i])])
            i = j
            continue
        j = i
        while j < n and j - i < 128 and not (j + 2 < n and vals[j] == vals[j + 1] == vals[j + 2]):
            j += 1
        out.append(j - i)
        out += bytes(int(v) for v in vals[i:j])
        i = j
    return bytes(out)


def encode(q: np.ndarray, style: str = "rle", header: bytes = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n") -> bytes:
    """RGBE texels (h, w, 4) -> .hdr bytes; style flat, rle or old."""
    q = np.asarray(q, np.uint8)
    h, w, _ = q.shape
    out = bytearray(header + b"\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        row = q[y]
        if style == "rle":
            assert 8 <= w <= 0x7FFF
            out += bytes([2, 2, w >> 8, w & 0xFF])
            for c in range(4):
                out += _rle_channel(row[:, c])
        elif style == "old":
            x = 0
            while x < w:
                out += row[x].tobytes()
                run = 0
                while x + 1 + run < w and (row[x + 1 + run] == row[x]).all() and run < 255:
                    run += 1
                if run >= 2:
                    out += bytes([1, 1, 1, run])
                    x += run
                x += 1
        else:
            out += row.tobytes()
    return bytes(out)
