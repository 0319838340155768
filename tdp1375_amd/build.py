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

SOURCES = ["tdp_search.hip", "tdp_filter.hip", "tdp_learn.hip", "tdp_canon.hip", "tdp_engine.hip", "jni_shim.cpp"]

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


def _compile(src: str, out_dir: str = OUT_DIR, defines=()) -> str:
    obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    extra = [f"-D{d}" for d in defines]
    if src.endswith(".hip"):
        cmd = [hipcc()] + COMMON + HIP_FLAGS + extra + ["-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = [hipcc()] + COMMON + extra + ["-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compiles and links the library.  `defines` adds -D macros (e.g.
    TDP_PHASE_TIMING for a profiling build) and `out` names the .so."""
    out_dir = OUT_DIR if not defines else OUT_DIR + "_" + "_".join(d.lower() for d in defines)
    os.makedirs(out_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, out_dir, defines), SOURCES))
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--phase-timing", action="store_true", help="profiling build with per-phase timers")
    ap.add_argument("--out", default=LIB)
    a = ap.parse_args()
    build(verbose=True, out=a.out, defines=("TDP_PHASE_TIMING",) if a.phase_timing else ())
    sys.exit(0)
