"""Build recipe of the native libraries (in-tree, gfx950).

  libfvm_gpu.so   HIP kernels + C ABI (include/fvm_gpu.h), hipcc --offload-arch=gfx950
  libfvm_mesh.so  host Voronoi builder and Rand48 (g++, OpenMP)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")

HIP_SOURCES = ["fvm_gpu.hip", "fvm_scan.hip", "fvm_fast.hip", "fvm_fused_f64.hip", "fvm_fused_f32.hip",
               "fvm_surrogate.hip", "fvm_diag.hip", "fvm_substep.hip"]
# -ffp-contract=off keeps every FP64 expression as written (no FMA contraction),
# which the reference-order and plain FP64 paths need to reproduce the
# reference's arithmetic.  The FP32 face path (its own translation unit)
# contracts to fused multiply-add.  packed-fp32-ops is disabled for all HIP code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
CONTRACT = {"fvm_fused_f32.hip": "-ffp-contract=fast"}
BUILD_DIR = os.path.join(CSRC, "_build")


def _run(cmd, cwd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_mesh():
    _run(["g++", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall", "-o",
          os.path.join(HERE, "libfvm_mesh.so"), "voronoi.cpp"], CSRC)


def build_gpu():
    """One object per source (each with its contraction mode), linked into one library;
    the sources compile in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = [os.path.join(BUILD_DIR, os.path.splitext(f)[0] + ".o") for f in HIP_SOURCES]

    def compile_one(i):
        f = HIP_SOURCES[i]
        _run(["hipcc", *HIP_FLAGS, CONTRACT.get(f, "-ffp-contract=off"), "-c", "-o", objs[i], f], CSRC)

    with ThreadPoolExecutor(max_workers=min(len(HIP_SOURCES), 4)) as ex:
        list(ex.map(compile_one, range(len(HIP_SOURCES))))
    out = os.path.join(HERE, "libfvm_gpu.so")
    tmp = out + ".tmp"
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], CSRC)
    os.replace(tmp, out)


def build_all():
    build_mesh()
    build_gpu()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "mesh": build_mesh, "gpu": build_gpu}[w]()
