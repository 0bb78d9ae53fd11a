"""Build the CPU oracle (test infrastructure only) with g++.

The oracle is the independent CPU restatement the HIP path is checked
against; it is never linked into libapt.  Outputs:

  oracle/liboracle.so         -O3, generic x86-64: what the tests load
  oracle/liboracle_x86v3.so   -O3 -march=x86-64-v3 (AVX2, FMA, BMI2): the
                              CPU baseline timed by bench.py on hosts that
                              have those instructions

-ffp-contract=off / no fast-math in both: the FP32 paths must round exactly
like the device code they are compared with, and the two builds give the
same bits (no fused multiply-adds are formed, reductions keep their order).

Usage: python oracle/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = sorted(HERE.glob("oracle_*.cpp"))
OUT = HERE / "liboracle.so"
OUT_V3 = HERE / "liboracle_x86v3.so"

FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
         "-fno-fast-math", "-Wall", "-Wno-unused-function"]
V3_FLAGS = ["-march=x86-64-v3", "-mtune=generic"]


def _link(out: Path, extra: list[str]) -> None:
    tmp = out.with_suffix(".so.tmp")
    cmd = ["g++", *FLAGS, *extra, *map(str, SOURCES), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)


def build(verbose: bool = True) -> Path:
    _link(OUT, [])
    _link(OUT_V3, V3_FLAGS)
    if verbose:
        print(f"[oracle] built {OUT} and {OUT_V3.name}", flush=True)
    return OUT


def baseline_flags() -> str:
    return " ".join(FLAGS[:1] + V3_FLAGS + ["-ffp-contract=off"])


if __name__ == "__main__":
    build()
    sys.exit(0)
