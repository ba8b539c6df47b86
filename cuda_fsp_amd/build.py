"""Builds the HIP library (gfx950) and the C++ facade in-tree.

    python -m cuda_fsp_amd.build          # or __graft_entry__.build()

Outputs (git-ignored, shipped with the tree):
    cuda_fsp_amd/lib/libfsp_hip.so   HIP kernels + C ABI (include/fsp_capi.h)
    cuda_fsp_amd/lib/libcufsp.so     C++ drop-in facade (cuFSP::, cme::) over the C ABI
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
CPP = os.path.join(HERE, "cpp")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIBDIR, "libfsp_hip.so")
CPP_LIB = os.path.join(LIBDIR, "libcufsp.so")

HIP_SOURCES = ["fsp_capi.hip", "fsp_model.hip", "fsp_mat.hip", "fsp_krylov.hip",
               "fsp_batch.hip", "fsp_batch_lockstep.hip", "fsp_batch_streams.hip", "fsp_bdf.hip",
               "fsp_adjoint.hip"]
CPP_SOURCES = ["cuFSP.cpp", "cme.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # every multiply/add separately rounded unless fma() is written
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-ffp-contract=off", "-fopenmp"]
CPP_INCLUDE = os.path.join(CPP, "include")


def compile_cpp_program(src: str, out: str) -> str:
    """Builds a C++ program against the facade (used by the C++ tests)."""
    _run([CXX, *CXX_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CPP_INCLUDE, src,
          "-o", out, "-L", LIBDIR, "-lcufsp", "-lfsp_hip", f"-Wl,-rpath,{LIBDIR}"])
    return out


def compile_hip_program(srcs: list, out: str) -> str:
    """Builds a HIP program against the facade with -fgpu-rdc (the device
    PropFun test: its __device__ propensity lives in another translation
    unit, as with the reference's separable compilation)."""
    _run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fgpu-rdc",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-I", CPP_INCLUDE, *srcs,
          "-o", out, "-L", LIBDIR, "-lcufsp", "-lfsp_hip", f"-Wl,-rpath,{LIBDIR}"])
    return out


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _hip_object(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
          os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        objs = list(ex.map(_hip_object, srcs))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_cpp() -> str | None:
    srcs = [os.path.join(CPP, "src", s) for s in CPP_SOURCES
            if os.path.exists(os.path.join(CPP, "src", s))]
    if not srcs:
        return None
    tmp = CPP_LIB + ".tmp"
    _run([CXX, *CXX_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(CPP, "include"), *srcs, "-o", tmp,
          "-L", LIBDIR, "-lfsp_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CPP_LIB)
    return CPP_LIB


def build_all(jobs: int = 8) -> list:
    out = [build_hip(jobs)]
    c = build_cpp()
    if c:
        out.append(c)
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
