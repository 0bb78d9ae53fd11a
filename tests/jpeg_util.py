"""Sample blocks for the JPEG encoder: 64x64 RGBX images chosen to reach every
part of baseline coding (flat blocks with only DC terms, smooth ramps, hard
edges and noise that fill all 63 AC terms with long codes, saturated colours
whose chroma sits at the ends of its range, 0xFF bytes in the entropy-coded
data)."""
from __future__ import annotations

import io

import numpy as np


def pillow_jpeg(rgbx, quality: int) -> bytes:
    """Pillow's (libjpeg-turbo) file for an RGBX block at 4:2:0."""
    from PIL import Image
    b = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(np.asarray(rgbx, np.uint8)[..., :3])).save(
        b, "JPEG", quality=int(quality), subsampling=2)
    return b.getvalue()


def sample_blocks(seed: int = 7) -> dict[str, np.ndarray]:
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:64, 0:64]
    out = {}

    def rgbx(r, g, b, xch=None):
        a = np.empty((64, 64, 4), np.uint8)
        a[..., 0], a[..., 1], a[..., 2] = r, g, b
        a[..., 3] = rng.integers(0, 256, (64, 64)) if xch is None else xch
        return a

    out["noise"] = rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)
    out["black"] = rgbx(0, 0, 0)
    out["white"] = rgbx(255, 255, 255)
    out["flat"] = rgbx(200, 120, 40)
    for i, c in enumerate([(255, 0, 0), (0, 255, 0), (0, 0, 255), (255, 255, 0), (0, 255, 255), (255, 0, 255)]):
        out[f"primary{i}"] = rgbx(*c)
    out["ramp"] = rgbx(x * 4, y * 4, 255 - (x + y) * 2)
    out["checker1"] = rgbx(*([((x + y) & 1) * 255] * 3))
    out["checker8"] = rgbx(((x // 8 + y // 8) & 1) * 255, 0, ((x // 8 + y // 8) & 1) * 255)
    out["stripes"] = rgbx((x % 3 == 0) * 255, (y % 5 == 0) * 255, ((x * y) % 7 == 0) * 255)
    out["edge"] = rgbx(np.where(x > 31, 255, 0), np.where(y > 20, 230, 10), 128)
    smooth = 128 + 100 * np.sin(x / 9.0) * np.cos(y / 7.0)
    out["smooth_noise"] = rgbx(np.clip(smooth + rng.normal(0, 6, (64, 64)), 0, 255).astype(np.uint8),
                               np.clip(smooth * 0.7 + rng.normal(0, 6, (64, 64)), 0, 255).astype(np.uint8),
                               np.clip(255 - smooth + rng.normal(0, 6, (64, 64)), 0, 255).astype(np.uint8))
    out["sparse"] = rgbx(0, 0, 0)
    out["sparse"][rng.integers(0, 64, 12), rng.integers(0, 64, 12), :3] = 255
    out["extremes"] = rgbx(*(rng.integers(0, 2, (3, 64, 64)) * 255))
    return out


def random_blocks(n: int, seed: int) -> np.ndarray:
    """n blocks mixing noise, smooth fields and edges."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:64, 0:64].astype(np.float32)
    blocks = np.empty((n, 64, 64, 4), np.uint8)
    kind = rng.integers(0, 3, n)
    for i in range(n):
        if kind[i] == 0:
            blocks[i] = rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)
        else:
            fx, fy, ph = rng.uniform(0.02, 0.5, 3)
            base = 128 + 120 * np.sin(fx * x + ph) * np.cos(fy * y)
            if kind[i] == 2:
                base = np.where(x + y * rng.uniform(-1, 1) > 32, 255 - base, base)
            for c in range(3):
                blocks[i, ..., c] = np.clip(base * rng.uniform(0.3, 1.0) + rng.normal(0, 3, (64, 64)), 0, 255)
            blocks[i, ..., 3] = 255
    return blocks
