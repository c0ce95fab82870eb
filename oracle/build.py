"""Build recipe of the CPU oracle library ``oracle/_build/liboracle.so`` (test infrastructure).

Compiled with g++ and the same floating-point contract as the kernels (no contraction, no
fast-math), OpenMP for the CPU-baseline timing. The target is x86-64-v3 (AVX2 + FMA) so the same
binary runs on the host of the GPU machine.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
OUT = HERE / "_build" / "liboracle.so"


def build(verbose: bool = True) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [
        "g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
        "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v3",
        f"-I{ROOT / 'include'}", str(HERE / "oracle.cpp"), "-o", str(tmp),
    ]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"[oracle] built {OUT}", flush=True)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
