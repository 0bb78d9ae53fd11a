"""Build the CPU oracle (test infrastructure only) with g++.

The oracle is the independent CPU restatement the HIP path is checked
against; it is never linked into libapt.  Output: oracle/liboracle.so.

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

# -ffp-contract=off / no fast-math: the FP32 paths must round exactly like the
# device code they are compared with.  -march is left generic so the library
# built here also runs on another x86-64 host.
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
         "-fno-fast-math", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = True) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["g++", *FLAGS, *map(str, SOURCES), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"[oracle] built {OUT}", flush=True)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
