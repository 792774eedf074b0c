"""Builds the HIP library ``sph07_amd/libsph07_amd.so`` for gfx950.

hipcc cross-compiles on a machine without a GPU.  The FP64 kernels are
compiled with ``-ffp-contract=off`` (no FMA formed: the reference's
operation order is kept); packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsph07_amd.so")
ARCH = os.environ.get("SPH07_ARCH", "gfx950")

# (source, floating-point contraction): the FP64 arm and the host code keep
# the reference's operation order; the FP32-compensated sweeps fuse the
# multiply-adds written within one expression ("on": decided per source
# expression, so every launch shape of a sweep computes the same bits; "fast"
# lets the backend fuse across statements, differently per instantiation).
# TwoSum itself is additions only, so contraction cannot touch it.
SOURCES = [("sph_kernels.hip", "off"), ("sph_fp32c.hip", "on"), ("sph_capi.cpp", "off"),
           ("mcm_reader.cpp", "off"), ("sph_output.cpp", "off")]
EXE = os.path.join(HERE, "sph07_run")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compiles every source to an object, then links the shared library."""
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    objs = []
    for src, contract in SOURCES:
        obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *COMMON, f"-ffp-contract={contract}", "-x", "hip", "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    # command-line driver, linked against the library next to it
    main_obj = os.path.join(out_dir, "sph07_main.o")
    cmd = [hipcc(), "-O2", "-std=c++17", "-c", os.path.join(CSRC, "sph07_main.cpp"), "-o", main_obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on sph07_main.cpp")
    cmd = [hipcc(), main_obj, "-o", EXE + ".tmp", f"-L{HERE}", "-lsph07_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of sph07_run failed")
    os.replace(EXE + ".tmp", EXE)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
