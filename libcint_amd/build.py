"""Build libmcint.so (HIP, gfx950) in-tree with hipcc.

    python -m libcint_amd.build [--jobs N]

Each translation unit is compiled separately (in parallel) and linked into
libcint_amd/libmcint.so.  Packed FP32 instructions are disabled for all
device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libmcint.so")

# longest compiles first
SOURCES = ([f"ip1_{g}.hip" for g in (14, *range(14))]
           + [f"chunk3_{g}.hip" for g in range(4)] + [f"fast3_{g}.hip" for g in range(4)]
           + [f"chunk_{g}.hip" for g in (17, 11, 18, 19, 20, 21, 12, 13, 14, 15, 16, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0)]
           + [f"fast_{g}.hip" for g in (7, 3, 8, 5, 6, 4, 2, 1, 0)] + ["eri.hip", "jk.hip", "api.hip"])

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCINT_ARCH", "gfx950")

CFLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
    "-Wno-unused-result",
]
if os.environ.get("MCINT_RESOURCE_USAGE"):
    CFLAGS += ["-Rpass-analysis=kernel-resource-usage"]


def _headers_mtime() -> float:
    names = [n for n in os.listdir(CSRC) if n.endswith((".h", ".inc"))]
    names = [os.path.join(CSRC, n) for n in names]
    names.append(os.path.join(HERE, "..", "include", "mcint.h"))
    return max(os.path.getmtime(n) for n in names)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if incremental and os.path.exists(obj):
        newest = max(_headers_mtime(), os.path.getmtime(os.path.join(CSRC, src)))
        if os.path.getmtime(obj) > newest:
            return obj
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    print(f"  compiled {src} in {time.time() - t:.1f}s", file=sys.stderr)
    if os.environ.get("MCINT_RESOURCE_USAGE"):
        with open(obj + ".remarks.txt", "w") as fh:
            fh.write(r.stderr)
    return obj


def build(jobs: int = 8, incremental: bool = False) -> str:
    """Compile every translation unit and link; incremental=True reuses
    objects newer than their source and every header."""
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found")
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args()
    t0 = time.time()
    print(build(a.jobs, a.incremental), f"({time.time() - t0:.1f}s)")
