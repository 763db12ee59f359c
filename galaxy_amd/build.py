"""Build the HIP library libgalaxy_lbvh.so in-tree for gfx950 (MI355X).

Usage: python -m galaxy_amd.build   (or galaxy_amd.build.build())
hipcc cross-compiles without a GPU, so this runs on any host with ROCm.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgalaxy_lbvh.so")
SOURCES = [
    "glx_kernels.hip", "glx_runtime.hip", "glx_api.hip", "glx_pool.hip", "glx_ic.hip", "glx_noalloc.hip",
    "glx_snapshot.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    # no packed-FP32 (v_pk_*_f32) instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(HERE, "..", "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgalaxy_lbvh.so")
    errs = [ln for ln in (r.stdout + r.stderr).splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln]
    if verbose and errs:
        sys.stderr.write("\n".join(errs) + "\n")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
