"""Builds libflybrain.so (all HIP sources, gfx950) and the N-API addon in-tree.

    python fly-brain_amd/build.py            # library only
    python fly-brain_amd/build.py --addon    # also the Node addon

hipcc compiles for gfx950 without a GPU present.  Floating point is built with
-ffp-contract=off: the reference kernel's arithmetic has no fused multiply-add
(it runs as wasm32), and bit-identical rasters need the same roundings.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libflybrain.so")
ADDON = os.path.join(LIBDIR, "flybrain_napi.node")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def sources():
    return sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))


def build_lib(verbose: bool = True) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in sources():
        obj = os.path.join(LIBDIR, os.path.basename(src) + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_addon(verbose: bool = True) -> str | None:
    """The N-API addon links libflybrain.so by rpath ($ORIGIN)."""
    hdr = "/usr/include/node"
    if not os.path.exists(os.path.join(hdr, "node_api.h")):
        if verbose:
            print("node headers absent: addon skipped", flush=True)
        return None
    src = os.path.join(PKG, "js", "fb_napi.cc")
    if not os.path.exists(src):
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", hdr, "-I", os.path.join(ROOT, "include"),
           src, "-L", LIBDIR, "-lflybrain", "-Wl,-rpath,$ORIGIN", "-o", ADDON]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return ADDON


if __name__ == "__main__":
    build_lib()
    if "--addon" in sys.argv:
        build_addon()
