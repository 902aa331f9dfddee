"""Builds the library for gfx950 with hipcc, in-tree:

    voronoids_amd/libvoronoids_hip.so   the shared library (ctypes, the C++ facade, the tests)
    voronoids_amd/libvoronoids_hip.a    the same objects as a static archive, for a caller that
                                        links statically (a Rust crate's build script, say; it
                                        also needs -lamdhip64 and the C++ runtime)

    python -m voronoids_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvoronoids_hip.so")
SRC = [os.path.join(HERE, "csrc", "runtime.hip"), os.path.join(HERE, "csrc", "engine.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # exact-arithmetic code: no fused multiply-add contraction anywhere
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd: list, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[0]))


def build(verbose: bool = False, out: str = LIB, defines: tuple = (), static: bool = True, tuned: bool = True) -> str:
    """compile the sources to objects (in parallel), link the shared library and, with static,
    archive the same objects next to it.  `defines` ("NAME=VALUE", ...) set round-kernel build
    constants; with tuned, the autotuner's stored constants (tuning.json) come first and `defines`
    override them."""
    stem = out[:-3] if out.endswith(".so") else out
    objs = [f"{stem}.{os.path.splitext(os.path.basename(s))[0]}.o" for s in SRC]
    if tuned:
        from .autotune import tuned_defines
        given = {d.split("=", 1)[0] for d in defines}
        defines = tuple(d for d in tuned_defines() if d.split("=", 1)[0] not in given) + tuple(defines)
    dflags = [f"-D{d}" for d in defines]
    with ThreadPoolExecutor(max_workers=len(SRC)) as ex:
        list(ex.map(lambda so: _run([HIPCC, *FLAGS, *dflags, "-c", "-o", so[1], so[0]], verbose), zip(SRC, objs)))
    tmp = out + ".tmp"
    _run([HIPCC, *FLAGS, "-shared", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    if static:
        a = stem + ".a"
        if os.path.exists(a):
            os.remove(a)
        _run(["ar", "rcs", a, *objs], verbose)
    for o in objs:
        os.remove(o)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
