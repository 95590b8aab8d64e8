"""Builds libopenvkl_hip.so (the HIP device's C-ABI library) in-tree.

hipcc compiles the HIP kernels for gfx950 and the C++ object layer; the
shared library lands in openvkl_amd/lib/ next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libopenvkl_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OPENVKL_HIP_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          "-fvisibility=hidden", "-Wall", "-Wno-unused-function"]
# Sampling / traversal numerics are binary32 without FMA contraction; packed
# FP32 instructions are disabled.
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
CXX_FLAGS = ["-x", "c++", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-isystem", "/opt/rocm/include"]


def _sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    flags = HIP_FLAGS if src.endswith(".hip") else CXX_FLAGS
    cmd = [HIPCC] + flags + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None) -> str:
    """Compile every source and link the library. Always rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    # The hit-iteration units take minutes each; start them first so they
    # overlap the rest instead of trailing it.
    order = sorted(srcs, key=lambda p: 0 if "hit" in os.path.basename(p) else 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        done = dict(zip(order, ex.map(_compile, order)))
    objs = [done[p] for p in srcs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
    sys.exit(0)
