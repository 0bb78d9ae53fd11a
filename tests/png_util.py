"""PNG files for the tests, independent of libapt's decoder: an encoder for
every colour type and bit depth (each row's filter cycles through the five
filter types; optional Adam7 interlacing) and a decoder built on Python's
zlib, both returning RGBA8 as a (h, w) uint32 array (R in the low byte)."""
from __future__ import annotations

import struct
import zlib

import numpy as np

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(t: bytes, d: bytes) -> bytes:
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def _pack_row(samples: np.ndarray, depth: int) -> bytes:
    """One row of samples (uint) at `depth` bits, big-endian, MSB first."""
    if depth == 16:
        return samples.astype(">u2").tobytes()
    if depth == 8:
        return samples.astype(np.uint8).tobytes()
    per = 8 // depth
    s = np.zeros(((len(samples) + per - 1) // per) * per, np.uint8)
    s[:len(samples)] = samples
    s = s.reshape(-1, per)
    out = np.zeros(s.shape[0], np.uint8)
    for k in range(per):
        out |= (s[:, k] << (8 - depth * (k + 1))).astype(np.uint8)
    return out.tobytes()


def _filter(rows, bpp):
    """Filter each row with type (row index mod 5)."""
    out = bytearray()
    prev = np.zeros(len(rows[0]) if rows else 0, np.int32)
    for y, r in enumerate(rows):
        cur = np.frombuffer(r, np.uint8).astype(np.int32)
        f = y % 5
        a = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]]) if len(cur) > bpp else np.zeros_like(cur)
        c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]]) if len(prev) > bpp else np.zeros_like(prev)
        b = prev
        if f == 0:
            v = cur
        elif f == 1:
            v = cur - a
        elif f == 2:
            v = cur - b
        elif f == 3:
            v = cur - ((a + b) >> 1)
        else:
            p = a + b - c
            pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
            v = cur - pred
        out.append(f)
        out.extend((v & 0xFF).astype(np.uint8).tobytes())
        prev = cur
    return bytes(out)


def encode(samples: np.ndarray, ctype: int, depth: int, interlace: bool = False, palette=None, trns=None) -> bytes:
    """samples: (h, w, channels) unsigned ints of the colour type's channels."""
    samples = np.asarray(samples)
    if samples.ndim == 2:
        samples = samples[:, :, None]
    h, w, ch = samples.shape
    assert ch == CHANNELS[ctype]
    bpp = max(1, ch * depth // 8)
    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        rows = [_pack_row(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
        raw += _filter(rows, bpp)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    if palette is not None:
        out += _chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    comp = zlib.compress(raw, 6)
    out += _chunk(b"IDAT", comp[: len(comp) // 2]) + _chunk(b"IDAT", comp[len(comp) // 2:])
    return out + _chunk(b"IEND", b"")


def expected_rgba(samples, ctype, depth, palette=None, trns=None) -> np.ndarray:
    """RGBA8 of raw samples: 16 bits -> high byte, 1-4 bit grey scaled to 8 bits,
    palette lookup, tRNS transparency; (h, w) uint32."""
    s = np.asarray(samples).astype(np.int64)
    if s.ndim == 2:
        s = s[:, :, None]
    h, w, _ = s.shape
    a = np.full((h, w), 255, np.int64)
    if ctype == 3:
        pal = np.asarray(palette, np.int64).reshape(-1, 3)
        idx = s[:, :, 0]
        r, g, b = pal[idx, 0], pal[idx, 1], pal[idx, 2]
        if trns is not None:
            t = np.full(256, 255, np.int64)
            t[:len(trns)] = list(trns)
            a = t[idx]
    else:
        def to8(x):
            return x >> 8 if depth == 16 else x if depth == 8 else x * 255 // ((1 << depth) - 1)
        if ctype in (0, 4):
            r = g = b = to8(s[:, :, 0])
            if ctype == 4:
                a = to8(s[:, :, 1])
            elif trns is not None:
                a = np.where(s[:, :, 0] == (trns[0] << 8 | trns[1]), 0, 255)
        else:
            r, g, b = to8(s[:, :, 0]), to8(s[:, :, 1]), to8(s[:, :, 2])
            if ctype == 6:
                a = to8(s[:, :, 3])
            elif trns is not None:
                key = [trns[2 * c] << 8 | trns[2 * c + 1] for c in range(3)]
                hit = (s[:, :, 0] == key[0]) & (s[:, :, 1] == key[1]) & (s[:, :, 2] == key[2])
                a = np.where(hit, 0, 255)
    return (r | (g << 8) | (b << 16) | (a << 24)).astype(np.uint32)


def decode(data: bytes) -> np.ndarray:
    """PNG (non-interlaced, 8-bit RGB / RGBA / grey / grey-alpha, or palette) -> RGBA8 with zlib + numpy."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    off, idat, plte, trns = 8, b"", None, None
    while off < len(data):
        n = struct.unpack(">I", data[off:off + 4])[0]
        t, d = data[off + 4:off + 8], data[off + 8:off + 8 + n]
        if t == b"IHDR":
            w, h, depth, ctype, _, _, lace = struct.unpack(">IIBBBBB", d)
        elif t == b"PLTE":
            plte = d
        elif t == b"tRNS":
            trns = d
        elif t == b"IDAT":
            idat += d
        off += 12 + n
    assert depth == 8 and lace == 0, (depth, lace)
    ch = CHANNELS[ctype]
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * ch)
    rows = np.zeros((h, w * ch), np.int32)
    prev = [0] * (w * ch)
    for y in range(h):
        f, cur = int(raw[y, 0]), raw[y, 1:].tolist()
        out = [0] * (w * ch)
        for x in range(w * ch):
            a = out[x - ch] if x >= ch else 0
            b = prev[x]
            c = prev[x - ch] if x >= ch else 0
            if f == 0:
                pr = 0
            elif f == 1:
                pr = a
            elif f == 2:
                pr = b
            elif f == 3:
                pr = (a + b) >> 1
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pr = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            out[x] = (cur[x] + pr) & 0xFF
        rows[y] = out
        prev = out
    s = rows.reshape(h, w, ch)
    return expected_rgba(s, ctype, 8, None if plte is None else np.frombuffer(plte, np.uint8),
                         None if trns is None else list(trns))
