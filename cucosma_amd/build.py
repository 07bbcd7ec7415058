"""Build libcosma_amd.so (the HIP kernels + C ABI) in-tree for gfx950.

Runs hipcc directly, one process per translation unit, then links the shared
library next to this file.  Run it by path, ``python cucosma_amd/build.py``,
so the package (which loads the library on import) is not imported first.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcosma_amd.so")

SOURCES = [
    "cosma_f32.hip",
    "cosma_f64.hip",
    "cosma_bf16x3.hip",
    "cosma_bf16x4.hip",
    "cosma_aux.hip",
    "cosma_api.cpp",
]

ARCH = "gfx950"
# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libcosma_amd.so")


def _compile(src: str, defines: list[str]) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *COMMON, *defines]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, fixed_shape: tuple[int, int, int] | None = None) -> str:
    """Compile every translation unit and link libcosma_amd.so; return its path.

    fixed_shape sets the (M, N, K) of the fixed-shape entry points
    cosma_sgemm / cosma_dgemm (default 1024 x 1024 x 1024).
    """
    os.makedirs(OBJDIR, exist_ok=True)
    defines = []
    if fixed_shape is not None:
        m, n, k = fixed_shape
        defines = [f"-DCOSMA_M={m}", f"-DCOSMA_N={n}", f"-DCOSMA_K={k}"]
    workers = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda s: _compile(s, defines), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
