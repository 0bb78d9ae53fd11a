"""Write tests/golden/cornell_conor.glb: the reference's
assets/CornellBoxConor.glb (the Cornell box its render host spawns,
apps/render-host/src/lib.rs:138) re-encoded as data the GPU tests can load
where the reference's assets are absent.  Geometry is flattened to world
space by the numpy reader (tests/gltf_util.read_glb), one mesh per material,
and the materials keep every factor and extension of the original; texture
references and images are dropped (the light's emission texture, a 956 KB
screenshot, would outweigh the rest of the fixture a hundred times).

Usage: python tests/golden/make_cornell_conor.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from tests.gltf_util import read_glb, write_glb  # noqa: E402

SRC = Path("/root/reference/assets/CornellBoxConor.glb")
OUT = Path(__file__).resolve().parent / "cornell_conor.glb"


def strip_textures(v):
    if isinstance(v, dict):
        return {k: strip_textures(x) for k, x in v.items() if not k.endswith("Texture")}
    if isinstance(v, list):
        return [strip_textures(x) for x in v]
    return v


def main() -> None:
    pos, nrm, mids, doc = read_glb(SRC)
    meshes, nodes = [], []
    for m in np.unique(mids):
        sel = mids == m
        P = pos[sel].reshape(-1, 3).astype(np.float32)
        N = nrm[sel].reshape(-1, 3).astype(np.float32)
        meshes.append([{"positions": P, "normals": N, "indices": np.arange(len(P), dtype=np.uint32),
                        "material": int(m)}])
        nodes.append({"mesh": len(meshes) - 1})
    mats = [strip_textures(m) for m in doc["materials"]]
    write_glb(OUT, nodes, meshes, mats, range(len(nodes)), doc.get("extensionsUsed", ()))
    print(f"wrote {OUT} ({OUT.stat().st_size} B, {len(pos)} triangles, {len(mats)} materials)")


if __name__ == "__main__":
    main()
