"""Writes cornell_fp64_4096px.npz: the FP64 oracle's converged radiance of the
Cornell box (512x512, depth 6, roulette from bounce 3) on a seeded subset of
4096 pixels, at 2^16 samples per pixel in 16 batches of 4096.

Arrays: pixels (the 4096 linear pixel indices, rng seed 0), mean (the 2^16-spp
FP64 mean per pixel, RGB), batch_var (the variance of one 4096-spp batch's
mean per pixel, from the 16 batches, ddof 1: the expected squared error of a
4096-spp estimate).  Batch b draws its samples with frame index 1 + b, so the
seeds (frame x 4096 + sample) never meet the 4096-spp frame 0 that the GPU
and the FP32 oracle render in the test.

Usage: python tests/golden/make_cornell_fp64.py   (about 90 s on 8 cores)
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from appearance_amd.scene import generate_scene  # noqa: E402
from tests._oracle import load_oracle, oracle_render  # noqa: E402

W = H = 512
SPP, BATCHES, DEPTH, RR = 4096, 16, 6, 3


def main():
    o = load_oracle()
    s = generate_scene("cornell")
    bvh = o.bvh(np.ascontiguousarray(s.positions, np.float32))
    pix = np.random.default_rng(0).choice(W * H, 4096, replace=False).astype(np.uint32)
    batches = []
    for b in range(BATCHES):
        m, _ = oracle_render(o, s, W, H, spp=SPP, max_bounces=DEPTH, rr_start=RR, pixels=pix,
                             precision=64, frame_idx=1 + b, bvh=bvh)
        batches.append(m.astype(np.float64))
        print(f"batch {b + 1}/{BATCHES}: mean {batches[-1].mean():.6f}", flush=True)
    bt = np.stack(batches)
    out = Path(__file__).with_name("cornell_fp64_4096px.npz")
    np.savez_compressed(out, pixels=pix, mean=bt.mean(0), batch_var=bt.var(0, ddof=1),
                        spp_total=np.array(SPP * BATCHES), batches=np.array(BATCHES))
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
