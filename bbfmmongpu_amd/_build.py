"""Builds the gfx950 shared libraries of the package (hipcc, in-tree).

libbbfmm_amd_s.so (real = float) and libbbfmm_amd_d.so (real = double) are
compiled from the same sources under csrc/, as the reference builds its
single- and double-precision executables from one tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 VALU instructions are disabled for every device compilation.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
          "-I", CSRC, "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def lib_path(prec: str) -> str:
    return os.path.join(PKG, f"libbbfmm_amd_{prec}.so")


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _compile(src: str, obj: str, defs):
    cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE_FLAGS, *defs, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-6000:]}")
    return obj


def build(precs=("s", "d"), jobs: int | None = None) -> list[str]:
    """Compile every HIP source for gfx950 and link one library per precision."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    out = []
    srcs = sources()
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = {}
        for prec in precs:
            defs = ["-DSINGLE"] if prec == "s" else []
            for s in srcs:
                obj = os.path.join(BUILD, f"{os.path.basename(s)}.{prec}.o")
                futs[ex.submit(_compile, s, obj, defs)] = prec
        objs = {p: [] for p in precs}
        for f in cf.as_completed(futs):
            objs[futs[f]].append(f.result())
    for prec in precs:
        lib = lib_path(prec)
        tmp = lib + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *sorted(objs[prec]), "-o", tmp,
               "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed for {lib}:\n{res.stderr[-4000:]}")
        os.replace(tmp, lib)
        out.append(lib)
    return out


if __name__ == "__main__":
    for p in build():
        print(p)
