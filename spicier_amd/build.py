"""Builds libspicier_hip.so in-tree with hipcc for gfx950.

Every HIP source is compiled with -ffp-contract=off (device models round as
the reference's CPU code; the LU kernels fuse explicitly with fma()) and with
packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libspicier_hip.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    """Every HIP source, the slowest to compile (the 64- and 32-unknown
    register-LU kernels) first so that the pool's last jobs are short."""
    def cost(f):
        heavy = ("newton" in f or "transient" in f) and ("_64" in f or "_32" in f)
        return (0 if heavy else 1, 0 if "_64" in f else 1, f)
    return [os.path.join(CSRC, f) for f in sorted(
        (f for f in os.listdir(CSRC) if f.endswith(".hip")), key=cost)]


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + [
        os.path.join(os.path.dirname(HERE), "include", "spicier_hip.h")]


def _compile(src, incremental=False):
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    if incremental and os.path.exists(obj):
        newest = max(os.path.getmtime(p) for p in [src, *_headers()])
        if os.path.getmtime(obj) >= newest:
            return obj
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose=False, incremental=False) -> str:
    """Compiles every HIP source (all of them unless `incremental`, which skips
    objects newer than their source and every header) and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    import sys
    build(verbose=True, incremental="--incremental" in sys.argv)
