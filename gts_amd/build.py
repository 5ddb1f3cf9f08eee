"""Builds the gfx950 library ``gts_amd/libgts.so`` and the ``gts_amd/bin/gts`` CLI.

Plain hipcc invocations (one object per translation unit, then one link), so
the build needs nothing beyond ROCm.  All device code is compiled for gfx950
only, with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgts.so")
BIN = os.path.join(HERE, "bin", "gts")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["gts_prims.hip", "gts_build.hip", "gts_search.hip", "gts_query.hip",
           "gts_api.hip", "gts_update.hip", "gts_learned.hip", "gts_mfma.hip",
           "gts_verify2.hip", "gts_cells.hip", "gts_lowdim.hip"]

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
         "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, msgs))
    return out, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    for _, msgs in results:
        if verbose and msgs.strip():
            print(msgs)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs,
                   check=True, capture_output=True)
    os.replace(tmp, LIB)
    cli = os.path.join(CSRC, "gts_cli.cpp")
    if os.path.exists(cli):
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"),
                        cli, "-o", BIN, "-L", HERE, "-lgts", "-Wl,-rpath,$ORIGIN/.."],
                       check=True, capture_output=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
