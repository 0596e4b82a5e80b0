"""Builds libfemocs_gpu.so in-tree with hipcc for gfx950.

    python -m femocs_amd.build            # compile (incremental on source mtimes)
    python -m femocs_amd.build --force    # recompile everything
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfemocs_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["fgpu_mesh.cpp", "fgpu_kernels.hip", "fgpu_api.hip", "fgpu_interp_host.cpp", "fgpu_interp.hip",
           "fgpu_tethex.hip", "fgpu_import.hip", "fgpu_q2.hip"]
# the transfer to atoms reproduces the reference's rounding: no floating-point contraction
EXTRA = {"fgpu_interp_host.cpp": ["-ffp-contract=off"], "fgpu_interp.hip": ["-ffp-contract=off"],
         "fgpu_tethex.hip": ["-ffp-contract=off"], "fgpu_import.hip": ["-ffp-contract=off"]}
HEADERS = ["fgpu_internal.h", "fgpu_kernels.h", "fgpu_interp.h", "fgpu_cg.inc", "fgpu_interp_build.inc", "fgpu_cheb.h"]

# -packed-fp32-ops: the compiler emits no packed-FP32 (v_pk_*_f32) instructions
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
         "-I" + os.path.join(ROOT, "include")]


def _stale(target: str, deps) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build(force: bool = False, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "femocs_gpu.h")]
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        if force or _stale(obj, [path] + hdrs):
            extra = EXTRA.get(src, [])
            cmd = [HIPCC] + FLAGS + extra + ["-c", path, "-o", obj]
            if src.endswith(".cpp"):
                cmd = [HIPCC, "-x", "hip"] + FLAGS + extra + ["-c", path, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
    if force or _stale(LIB, objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
