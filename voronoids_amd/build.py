"""Builds voronoids_amd/libvoronoids_hip.so with hipcc for gfx950 (in-tree).

    python -m voronoids_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvoronoids_hip.so")
SRC = [os.path.join(HERE, "csrc", "runtime.hip"), os.path.join(HERE, "csrc", "engine.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # exact-arithmetic code: no fused multiply-add contraction anywhere
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose: bool = False, out: str = LIB, defines: tuple = ()) -> str:
    """compile the library; `defines` ("NAME=VALUE", ...) select tuning variants"""
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-o", tmp, *SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
