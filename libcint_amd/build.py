"""Build libmcint.so (HIP, gfx950) in-tree with hipcc.

    python -m libcint_amd.build [--jobs N]

Each translation unit is compiled separately (in parallel) and linked into
libcint_amd/libmcint.so.  Packed FP32 instructions are disabled for all
device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import re
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
           + [f"fast_{g}.hip" for g in (7, 3, 8, 5, 6, 4, 2, 1, 0)] + ["eri.hip", "jk.hip", "api.hip", "noalloc.hip", "sr_roots.hip"])

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


def _closure(path: str, seen=None) -> list:
    """The file and every local header it includes (#include "..."), recursively."""
    seen = [] if seen is None else seen
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return seen
    seen.append(path)
    with open(path, errors="replace") as fh:
        for line in fh:
            m = re.match(r'\s*#\s*include\s+"([^"]+)"', line)
            if m:
                _closure(os.path.join(os.path.dirname(path), m.group(1)), seen)
    return seen


def _digest(src: str) -> str:
    """Hash of the flags and the text of a source and its local includes."""
    h = hashlib.sha256(" ".join(CFLAGS).encode())
    for f in _closure(os.path.join(CSRC, src)):
        with open(f, "rb") as fh:
            h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    stamp = obj + ".sha"
    digest = _digest(src)
    if incremental and os.path.exists(obj) and os.path.exists(stamp):
        with open(stamp) as fh:
            if fh.read().strip() == digest:
                return obj
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    with open(stamp, "w") as fh:
        fh.write(digest)
    print(f"  compiled {src} in {time.time() - t:.1f}s", file=sys.stderr)
    if os.environ.get("MCINT_RESOURCE_USAGE"):
        with open(obj + ".remarks.txt", "w") as fh:
            fh.write(r.stderr)
    return obj


def build(jobs: int = 8, incremental: bool = False) -> str:
    """Compile every translation unit and link; incremental=True reuses
    objects built from the same flags and the same text of the source and
    of every local header it includes."""
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
