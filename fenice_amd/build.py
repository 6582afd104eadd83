"""Build of the device library libfvg.so (HIP for gfx950) in-tree.

Flags that matter for the results: -ffp-contract=off and no fast-math keep every FP64 operation the
reference's (no fused multiply-add, IEEE division and square root); packed FP32 instructions are switched off
(-packed-fp32-ops). The library links only the HIP runtime.
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfvg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FVG_ARCH", "gfx950")

SOURCES = ["fvg_kernels_f64.hip", "fvg_kernels_f32.hip", "fvg_kernels_f4.hip", "fvg_kernels.hip", "fvg_api.hip", "fvg_batch.hip", "fvg_learned.hip", "fvg_noalloc.hip",
           "fvg_config.cpp", "fvg_checkpoint.cpp", "fvg_npz.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fno-gpu-rdc"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def build(verbose: bool = False) -> str:
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(obj_dir, src + ".o")
        if src.endswith(".hip"):
            cmds.append([HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", path, "-o", obj])
        else:
            cmds.append([HIPCC, *COMMON, "-c", path, "-o", obj])
        objs.append(obj)
        if verbose:
            print(" ".join(cmds[-1]), flush=True)
    with ThreadPoolExecutor(max_workers=min(6, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
