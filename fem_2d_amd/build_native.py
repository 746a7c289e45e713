"""Compile the native libraries in-tree.

* ``fem_2d_amd/libfem2d_amd.so``: host input builder (C++17, g++) and the HIP
  assembly kernels (hipcc, gfx950), linked into one shared library exposing
  the C ABI of ``include/fem2d.h``.
* ``oracle/_build/liboracle.so``: the CPU restatement of the reference
  algorithm (test infrastructure; the product library never loads it).

Run ``python -m fem_2d_amd.build_native`` or call :func:`build`.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfem2d_amd.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

HOST_SOURCES = ["host/mesh.cpp", "host/domain.cpp", "host/capi_host.cpp"]
HIP_SOURCES = ["assemble.hip", "domain_build.hip", "petsc.hip", "eigen.hip", "sparse_eigen.hip"]

CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra"]
HIPFLAGS = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "--offload-arch=gfx950",
    # the reference evaluates every product and sum separately; keep the
    # compiler from fusing them so the shape tables match it operation by operation
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return res


def build_library(verbose: bool = False) -> str:
    """Compile and link libfem2d_amd.so (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src).replace(".cpp", ".o"))
        _run(["g++", *CXXFLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    hipcc = _hipcc()
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
        _run([hipcc, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, f"-L{rocm_lib}",
          f"-Wl,-rpath,{rocm_lib}", "-lrocsolver", "-lrocblas"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    """Compile the oracle (CPU restatement of the reference) with g++."""
    os.makedirs(os.path.dirname(ORACLE_LIB), exist_ok=True)
    tmp = ORACLE_LIB + ".tmp"
    _run(
        ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", os.path.join(ORACLE_DIR, "oracle.cpp"), "-o", tmp,
         "-lpthread"],
        verbose,
    )
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB, "and", ORACLE_LIB)
