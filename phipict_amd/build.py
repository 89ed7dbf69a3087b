"""Build the gfx950 library ``phipict_amd/_lib/libphipict_hip.so`` from ``phipict_amd/csrc``.

The HIP sources are compiled with hipcc for gfx950 only (packed-FP32 instructions disabled),
the host-only AMG setup with g++, and everything is linked into one shared library whose
C ABI is declared in ``include/phipict_hip.h``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libphipict_hip.so"
BUILD = ROOT / "_build" / "phipict_amd"

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"-I{ROOT / 'include'}"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... (exit {res.returncode})")


def build(verbose: bool = False) -> Path:
    """Compile every source and link the library. Always rebuilds."""
    hipcc = _hipcc()
    BUILD.mkdir(parents=True, exist_ok=True)
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    jobs = []
    objs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = BUILD / (src.stem + ".o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    for src in sorted(CSRC.glob("*.cpp")):
        obj = BUILD / (src.stem + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "4"))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
