"""Builds the library's batched solver (fsp_batch.hip, fsp_batch_lockstep.hip,
fsp_model.hip) for the CPU with g++ against the SIMT stand-in header in this folder, as
tests/simt/_build/libfsp_batch_simt.so.  The sources are compiled unchanged:
the stand-in supplies the HIP runtime, kernel launches, barriers and shuffles.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "cuda_fsp_amd", "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libfsp_batch_simt.so")

SOURCES = [os.path.join(CSRC, "fsp_batch.hip"), os.path.join(CSRC, "fsp_batch_lockstep.hip"),
           os.path.join(CSRC, "fsp_model.hip"), os.path.join(HERE, "simt_support.cpp")]
DEPS = SOURCES + [os.path.join(HERE, "hip", "hip_runtime.h")] + [
    os.path.join(CSRC, h) for h in ("fsp_batch_impl.h", "fsp_expm.h", "fsp_formats.h",
                                    "fsp_device.h", "fsp_internal.h", "fsp_krylov_dev.h")]


def build(force: bool = False) -> str:
    """Compiles the CPU build when a source is newer than the library."""
    if (not force and os.path.exists(LIB)
            and os.path.getmtime(LIB) >= max(os.path.getmtime(p) for p in DEPS)):
        return LIB
    os.makedirs(OUT_DIR, exist_ok=True)
    # a file of this process's own: test workers that rebuild at the same time
    # each install a whole library (the rename is atomic)
    tmp = f"{LIB}.{os.getpid()}.tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-g", "-fPIC", "-shared",
           # fibers switch stacks with _setjmp/_longjmp
           "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0",
           "-ffp-contract=off", "-Wno-unused-function", "-Wno-unused-variable",
           "-I", HERE, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           "-x", "c++", *SOURCES, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("simt build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB
