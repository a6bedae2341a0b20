"""Build the gfx950 stage library in-tree (sdbe_amd/libsdbe.so).

Run ``python -m sdbe_amd.build`` or call :func:`build_all`.  hipcc cross-
compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsdbe.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 ops stay off (v_pk_*_f32 are not emitted)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass warns that it does not know the packed-fp32 feature: harmless
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_stage_lib():
    tmp = LIB + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "sdbe_stage.hip")])
    os.replace(tmp, LIB)
    return LIB


def build_all(verbose=False):
    built = [build_stage_lib()]
    if verbose:
        for b in built:
            print("built", os.path.relpath(b, ROOT))
    return built


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
