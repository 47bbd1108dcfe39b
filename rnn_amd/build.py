"""Build libgsknn.so for gfx950 in-tree (rnn_amd/lib/).

    python -m rnn_amd.build

Compiles the HIP kernels and the C-ABI layer with hipcc into one shared
library.  Device code is built for gfx950 only, with contraction of a*b + c
disabled (the exact kernels reproduce a rounded product followed by a rounded
sum) and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libgsknn.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = ["gsknn_kernels.hip", "gsknn_ring.hip", "gsknn_rerank.hip", "gsknn_f32.hip", "gsknn_sweep.hip",
           "gsknn_l1.hip", "gsknn_noalloc.hip", "gsknn_api.cpp"]

HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose: bool = False) -> str:
    """Compile every source (in parallel, at most MAX_JOBS or 8 at once) and
    link the library; raises CalledProcessError on the first failure."""
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    cmds, objs = [], []
    for src in sources():
        obj = os.path.join(LIBDIR, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([hipcc] + HIPFLAGS + lang + ["-c", src, "-o", obj])
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(compile_one, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
