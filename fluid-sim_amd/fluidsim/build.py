"""Build libfluidsim.so (HIP, gfx950) in-tree with hipcc.

The library is the product: kernels plus the C ABI of include/fluidsim.h.
hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # fluid-sim_amd/
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libfluidsim.so")
ARCH = os.environ.get("FLUIDSIM_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # identical bits for identical formulas across kernels; FMAs are written out
    "-ffp-contract=off",
    # no packed-FP32 VALU instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    return [os.path.join(CSRC, "fs_solver.hip")]


NODE_DIR = os.path.join(PKG, "node")
NODE_ADDON = os.path.join(NODE_DIR, "fluidsim.node")
NODE_INCLUDE = "/usr/include/node"


def build_node(verbose: bool = False):
    """Build the Node-API addon (fluidsim.node) over libfluidsim.so with gcc.

    Returns the addon path, or None when the Node-API headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    tmp = NODE_ADDON + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-I", NODE_INCLUDE, "-I", os.path.join(ROOT, "include"),
           os.path.join(NODE_DIR, "fluidsim_napi.c"), "-o", tmp, "-L", LIBDIR, "-lfluidsim",
           "-Wl,-rpath,$ORIGIN/../lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, NODE_ADDON)
    return NODE_ADDON


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["/opt/rocm/bin/hipcc", *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *sources(), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_node(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
