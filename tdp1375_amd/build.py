"""Builds libtdp1375_jni.so (HIP kernels + host engine + C ABI + JNI natives).

    python -m tdp1375_amd.build

hipcc cross-compiles for gfx950 without a GPU.  Objects go to
tdp1375_amd/_build/, the library to tdp1375_amd/libtdp1375_jni.so (in-tree, so
it travels with the source tree and is what tests, smoke and bench load).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtdp1375_jni.so")
ARCH = os.environ.get("TDP_OFFLOAD_ARCH", "gfx950")

SOURCES = ["tdp_search.hip", "tdp_filter.hip", "tdp_engine.hip", "jni_shim.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-fvisibility=hidden",
          "-I", os.path.join(HERE, "..", "include")]
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=off"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [hipcc()] + COMMON + HIP_FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = [hipcc()] + COMMON + ["-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
