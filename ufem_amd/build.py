"""Build libufem_gpu.so in-tree: hand-written HIP kernels for gfx950 (hipcc)
plus the host C layer (gcc), linked into ufem_amd/libufem_gpu.so.

    python -m ufem_amd.build          (or __graft_entry__.build())
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libufem_gpu.so")

HIP_SOURCES = ["ugk_kernels.hip", "ugk_unet.hip", "ugk_noalloc.hip"]
C_SOURCES = ["ufem_sys.c", "ufem_solve.c", "ufem_ssor.c", "ufem_coarse.c", "fem_dropin.c", "fem_fillk.c", "ufem_batch.c", "ufem_net.c",
             "ufem_input.c", "ufem_monte.c"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
             # keep packed FP32 instructions out of the generated code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
C_FLAGS = ["-O2", "-std=c99", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + "\n".join(lines))
    return "\n".join(lines)


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        msg = _run([HIPCC, *HIP_FLAGS, "-I", CSRC, "-I", INC, "-c", os.path.join(CSRC, src), "-o", o])
        if verbose and msg:
            print(msg)
        objs.append(o)
    cc = shutil.which("gcc") or "cc"
    for src in C_SOURCES:
        o = os.path.join(OBJ, src.replace(".c", ".o"))
        msg = _run([cc, *C_FLAGS, "-I", CSRC, "-I", INC, "-c", os.path.join(CSRC, src), "-o", o])
        if verbose and msg:
            print(msg)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lm", "-lpthread", "-ldl"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
