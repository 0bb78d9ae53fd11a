"""Build libapt.so (HIP for gfx950 + C++ host runtime) in-tree with hipcc.

Usage: python -m appearance_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libapt.so"
OBJ_DIR = PKG / "_build"

SOURCES = [
    "api_core.cpp",
    "bvh_host.cpp",
    "scenes.cpp",
    "trace.hip",
    "renderer.cpp",
    "scene_gpu.cpp",
    "integrator.hip",
    "server.cpp",
    "server.hip",
    "tuner.cpp",
    "restir.hip",
    "restir_two_level.hip",
    "restir_textured.hip",
    "restir_textured_two_level.hip",
    "bvh_gpu.hip",
    "gltf.cpp",
    "image.cpp",
    "tlas.hip",
    "jpeg.hip",
    "jpeg_decode.cpp",
]

ARCH = os.environ.get("APT_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off on host and device: the triangle test and the shading
# maths must round identically on both sides.  Packed FP32 instructions are
# disabled for all device code.
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(PKG.parent / "include"),
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: Path, obj: Path) -> None:
    cmd = [_hipcc(), "-x", "hip", *COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every source and link libapt.so; always rebuilds."""
    OBJ_DIR.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    objs = [OBJ_DIR / (s.name + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda p: _compile(*p), zip(srcs, objs)))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [_hipcc(), *COMMON, "-shared", "-pthread", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"[appearance_amd] built {OUT} from {len(srcs)} sources for {ARCH}", flush=True)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
