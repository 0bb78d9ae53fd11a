"""Writes tests/golden/bvh_stress_counts.json: per-ray visit counts of the
oracle's canonical BVH2 (binned SAH, 32 bins, leaf <= 4, 32 B nodes, 36 B
triangles) on the BVH-stress workload (1,048,576 triangles, seed 1; 1M
coherent + 1M incoherent rays, seed 2).

These counts fix the algorithmic bytes per ray used for the roofline:
closest = 48 + 32*N_box + 36*N_tri bytes, any = 36 + 32*N_box + 36*N_tri.
They do not depend on the GPU BVH width."""
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from appearance_amd.raytrace import rays_bvh_stress  # noqa: E402
from appearance_amd.scene import generate_scene  # noqa: E402
from tests._oracle import load_oracle  # noqa: E402


def main():
    o = load_oracle()
    s = generate_scene("bvh_stress", seed=1)
    rays = rays_bvh_stress(seed=2, side=1024, n_incoherent=1 << 20).view("f4").reshape(-1, 8)
    t0 = time.time()
    b = o.bvh(s.positions)
    out = {"n_tris": int(s.n_tris), "n_rays": int(rays.shape[0]), "oracle_nodes": int(o.lib.or_bvh_node_count(b.h)),
           "build_s": time.time() - t0}
    halves = {"coherent": rays[: 1 << 20], "incoherent": rays[1 << 20:], "all": rays}
    for q in ("closest", "any"):
        for name, r in halves.items():
            t0 = time.time()
            h = b.trace(r, any_hit=(q == "any"))
            nbox, ntri = b.last_counts
            n = r.shape[0]
            hits = int((h.view("u4")[:, 3] != 0xFFFFFFFF).sum()) if q == "closest" else int(h.sum())
            out[f"{q}_{name}"] = {"box_per_ray": nbox / n, "tri_per_ray": ntri / n, "hits": hits,
                                  "seconds": time.time() - t0}
    for q, base in (("closest", 48), ("any", 36)):
        c = out[f"{q}_all"]
        c["alg_bytes_per_ray"] = base + 32 * c["box_per_ray"] + 36 * c["tri_per_ray"]
    Path(__file__).with_name("bvh_stress_counts.json").write_text(json.dumps(out, indent=1))
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
