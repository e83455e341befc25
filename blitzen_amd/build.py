"""Builds libbzcull.so (HIP, gfx950) in-tree.

Every HIP source is compiled with -ffp-contract=off (the compiler fuses no multiply-add; the cull arithmetic fuses
only where it calls fma explicitly, as the oracle does, so device fp32 results equal a CPU build of the same
expressions), HIP's default correctly rounded fp32 division and square root, IEEE denormals, and packed FP32
instructions disabled. Objects are rebuilt every call; the link is one hipcc invocation.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libbzcull.so"

SOURCES = ["cull.hip", "pyramid.hip", "cluster.hip", "instanced.hip", "capi.cpp", "camera.cpp", "synth.cpp",
           "autotune.cpp", "server.cpp", "interop.cpp", "noalloc.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-result",
]


def kernel_sources_digest() -> str:
    """SHA-256 (first 16 hex digits) of the library's sources and public header: profiles record it so that a
    measurement can be told apart from one of other kernels."""
    import hashlib

    h = hashlib.sha256()
    files = sorted(p for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp", ".h")) + [PKG.parent / "include" /
                                                                                         "bzcull.h"]
    for p in files:
        h.update(p.name.encode())
        h.update(p.read_bytes())
    return h.hexdigest()[:16]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libbzcull)")


def _compile(src: str) -> Path:
    obj = BUILD / (src + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    present = [s for s in SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=min(8, len(present))) as ex:
        objs = list(ex.map(_compile, present))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
