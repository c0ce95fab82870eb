"""In-tree build of the gfx950 kernel library ``_lib/libpht_splat.so``.

Each ``csrc/*.hip`` file is compiled to an object with ``hipcc --offload-arch=gfx950`` (files in
parallel), then linked into one shared library exporting the C ABI declared in
``include/pht_splat.h``. Floating-point flags are part of the parity contract: no contraction
(``-ffp-contract=off``) and no fast-math, so the kernels compute the same bits as the CPU oracle.
Packed FP32 instructions are disabled for the device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
BUILD = PKG / "_build"
LIBNAME = "libpht_splat.so"
ARCH = os.environ.get("PHT_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build the kernels")
    return path


def _compile(src: Path) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every kernel source for gfx950 and link the shared library. Always rebuilds."""
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    LIBDIR.mkdir(parents=True, exist_ok=True)
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"[photara_studio_amd] built {out} ({out.stat().st_size // 1024} KiB)", flush=True)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
