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
HASHPIPE_LIB = os.path.join(HERE, "libsdbe_hashpipe.so")
INCLUDE = os.path.join(ROOT, "include")
COMPAT = os.path.join(CSRC, "hashpipe_compat")

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
    _run([_hipcc(), *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "sdbe_stage.hip"), os.path.join(CSRC, "sdbe_pfb.hip"),
          os.path.join(CSRC, "sdbe_formats.hip"), os.path.join(CSRC, "sdbe_qa.hip")])
    os.replace(tmp, LIB)
    return LIB


# Plugin builds: (library, SWARM_RATE_FRAC, band-trimmed resampler).  The
# default is the 39 -> 64 resampler at the 6/11 rate; the others are the
# reference thread's band-trimmed path at each SWARM rate preset (their
# databuf layouts follow the preset, include/aphids_databuf.h).
PLUGIN_VARIANTS = [
    (HASHPIPE_LIB, 6, 0),
    (os.path.join(HERE, "libsdbe_hashpipe_trim6.so"), 6, 1),
    (os.path.join(HERE, "libsdbe_hashpipe_trim8.so"), 8, 1),
    (os.path.join(HERE, "libsdbe_hashpipe_trim10.so"), 10, 1),
    (os.path.join(HERE, "libsdbe_hashpipe_trim11.so"), 11, 1),
]


def build_hashpipe_plugin(lib=HASHPIPE_LIB, rate=6, trimmed=0):
    """vdif_inout_gpu_thread plugin, the file-fed input thread and the
    file-writing output thread, and the host bookkeeping, against the
    in-process hashpipe API stand-in (production builds use the real hashpipe
    headers)."""
    tmp = lib + ".tmp"
    srcs = [os.path.join(CSRC, "vdif_inout_gpu_thread.c"), os.path.join(CSRC, "vdif_in_file_thread.c"),
            os.path.join(CSRC, "vdif_out_file_thread.c"), os.path.join(CSRC, "aphids_host.c"),
            os.path.join(CSRC, "aphids_databuf_hip.c"), os.path.join(COMPAT, "hashpipe_compat.c")]
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    # plain C against the HIP runtime API (the HIP headers need the platform macro);
    # -Bsymbolic: each variant's calls stay inside it when several are loaded
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall", "-D__HIP_PLATFORM_AMD__",
          "-DSWARM_RATE_FRAC=%d" % rate, "-DSDBE_PLUGIN_TRIMMED=%d" % trimmed,
          "-I", INCLUDE, "-I", COMPAT, "-I", os.path.join(rocm, "include"), "-o", tmp, *srcs,
          "-L", HERE, "-lsdbe", "-L", os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-Bsymbolic",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib"), "-lpthread", "-lm"])
    os.replace(tmp, lib)
    return lib


def build_all(verbose=False):
    built = [build_stage_lib()] + [build_hashpipe_plugin(*v) for v in PLUGIN_VARIANTS]
    if verbose:
        for b in built:
            print("built", os.path.relpath(b, ROOT))
    return built


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
