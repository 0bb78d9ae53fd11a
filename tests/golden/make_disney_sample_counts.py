"""Writes disney_sample_counts.json: per-sample work of the Disney frame
(256 UV spheres + ground, seed 3; 1920x1088, 4 spp, depth 8, roulette from
bounce 3, seed 0, frame 0) counted by the FP32 oracle path tracer on every
pixel of the frame, with the same per-pixel seeds as the device integrator:
extension (closest-hit) and shadow (any-hit) rays per sample, and the box and
triangle tests per ray of the oracle's canonical BVH2 (binned SAH, 32 bins,
leaf <= 4, 32 B per box, 36 B per triangle).

From them the algorithmic bytes:
  closest-hit ray  48 B (32 B ray in, 16 B hit out) + 32 N_box + 36 N_tri
  any-hit ray      36 B (32 B ray in, 4 B verdict out) + 32 N_box + 36 N_tri
  sample           sum of its rays + 256 B per path segment (one per
                   extension ray: hit 16, path state read and written 2 x 48,
                   three vertices 96, new ray 32, queue entries 16) + 16 B film

Usage: python tests/golden/make_disney_sample_counts.py
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from appearance_amd.scene import generate_scene  # noqa: E402
from tests._oracle import load_oracle, oracle_render  # noqa: E402

W, H, SPP, DEPTH, RR, SEED = 1920, 1088, 4, 8, 3, 0


def main():
    o = load_oracle()
    s = generate_scene("disney", seed=3)
    bvh = o.bvh(np.ascontiguousarray(s.positions, np.float32))
    t0 = time.time()
    _, c = oracle_render(o, s, W, H, spp=SPP, max_bounces=DEPTH, rr_start=RR, seed=SEED, bvh=bvh,
                         tally=True)
    dt = time.time() - t0
    n = W * H * SPP
    ext, sh = c["extension_rays"], c["shadow_rays"]
    nb_c, nt_c = c["closest_box_tests"] / ext, c["closest_tri_tests"] / ext
    nb_a, nt_a = c["any_box_tests"] / sh, c["any_tri_tests"] / sh
    b_ext = 48 + 32 * nb_c + 36 * nt_c
    b_sh = 36 + 32 * nb_a + 36 * nt_a
    seg = ext / n
    b_sample = ext / n * b_ext + sh / n * b_sh + 256 * seg + 16
    out = {
        "workload": "disney seed 3, 1920x1088, 4 spp, depth 8, rr from 3, seed 0, frame 0, all pixels",
        "samples": n,
        "counts": c,
        "per_sample": {"extension_rays": ext / n, "shadow_rays": sh / n, "segments": seg},
        "per_ray": {"closest_box_tests": nb_c, "closest_tri_tests": nt_c, "any_box_tests": nb_a,
                    "any_tri_tests": nt_a},
        "alg_bytes": {"closest_ray": b_ext, "any_ray": b_sh, "sample": b_sample,
                      "extend_stage_per_sample": ext / n * b_ext},
        "oracle_seconds": round(dt, 1),
    }
    path = Path(__file__).with_name("disney_sample_counts.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
