"""Builds lshongpu_amd/liblshgpu.so (hand-written HIP for gfx950) with hipcc.

Usage: python -m lshongpu_amd.build   (or lshongpu_amd.build.build())

Every call recompiles every source; objects go to lshongpu_amd/_build/.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblshgpu.so")

SOURCES = [
    "common.cpp",
    "api.cpp",
    "autotune.cpp",
    "serve.cpp",
    "engine.cpp",
    "select.hip",
    "scan_exact.hip",
    "scan_exact_mfma.hip",
    "scan_sketch.hip",
    "scan_mfma.hip",
    "cells.hip",
    "rerank.hip",
    "sketch.hip",
    "minhash.hip",
    "lsh.hip",
    "stats.hip",
    "util.hip",
    "noalloc.hip",
    "crosspoly_dense.hip",
    "simhash_mfma.hip",
    "io.cpp",
]

# Command-line runners linked against the library: the reference's knn main
# (kernel.cu:119-196) and a GPU truth-file generator (goldstandard.cpp).
PROGRAMS = [("knn_main.cpp", "knn"), ("gold_main.cpp", "goldstandard")]

ARCH = os.environ.get("LSHGPU_ARCH", "gfx950")

# -ffp-contract=off: every fused multiply-add the parity-critical code needs
# is written explicitly; nothing else may be contracted behind its back.
# -packed-fp32-ops: no packed FP32 VALU instructions are generated.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build liblshgpu.so)")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-x", "hip", "-c", *HIPFLAGS, "-o", obj, os.path.join(CSRC, src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    errs = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{errs}")
    if errs.strip():
        sys.stderr.write(errs + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    for main, exe in PROGRAMS:
        cmd = [hipcc(), "-O2", "-std=c++17", "-o", os.path.join(HERE, exe), os.path.join(CSRC, main),
               f"-L{HERE}", "-llshgpu", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {exe} failed:\n{r.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
