"""Build the gfx950 libraries in-tree.

    python -m eigen_amd.build            # compile everything (always rebuilds)

Produces, next to this file (or in $EIGEN_AMD_LIBDIR when set):
    _lib/libeigen_amd.so        the C-ABI (include/eigen_amd.h) and all kernels
    _lib/libeigen_amd_blas.so   sgemm_/dgemm_/xerbla_ (include/eigen_amd_blas.h)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.environ.get("EIGEN_AMD_OBJDIR") or os.path.join(HERE, "_build")
LIB = os.environ.get("EIGEN_AMD_LIBDIR") or os.path.join(HERE, "_lib")

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 instructions are disabled for all device code.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["gemm_f32.hip", "gemm_f64.hip", "gemm_split.hip", "gemv.hip", "level3_kernels.hip", "fill.hip", "grouped.hip",
               "mlp.hip", "complex_kernels.hip", "noalloc.hip", "level3.cpp", "complex_gemm.cpp",
               "capi.cpp", "planner.cpp", "tuner.cpp"]
BLAS_SOURCES = ["blas.cpp"]
# Per-file code generation choices.  The FP64 kernels keep their
# accumulators in the VGPR form of the MFMA: with AGPR accumulators the
# 128x128 tile copied every accumulator through VGPRs on each k slice.
FILE_FLAGS = {"gemm_f64.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
              # the grouped kernel's accumulators in VGPRs: 112 registers, four waves per SIMD
              "grouped.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    extra = FILE_FLAGS.get(src, []) + os.environ.get("EIGEN_AMD_HIPFLAGS", "").split()
    cmd = [HIPCC] + HIPFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    sources = [s for s in LIB_SOURCES + BLAS_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = dict(zip(sources, ex.map(_compile, sources)))
    lib_objs = [objs[s] for s in LIB_SOURCES if s in objs]
    libpath = os.path.join(LIB, "libeigen_amd.so")
    subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", libpath + ".tmp"] + lib_objs,
                   check=True)
    os.replace(libpath + ".tmp", libpath)
    blaspath = os.path.join(LIB, "libeigen_amd_blas.so")
    subprocess.run([HIPCC, "-shared", "-fPIC", "-o", blaspath + ".tmp", objs["blas.cpp"], "-L" + LIB,
                    "-leigen_amd", "-Wl,-rpath,$ORIGIN"], check=True)  # xerbla_ stays interposable
    os.replace(blaspath + ".tmp", blaspath)
    if verbose:
        print("built", libpath, blaspath)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
