"""Build the HIP library libbielcc_hip.so for gfx950 (in-tree, bielcc_amd/lib/).

Each .hip translation unit is compiled separately (in parallel) and linked
with hipcc; the library links rocBLAS for the LU's trailing-matrix ZGEMM.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libbielcc_hip.so")
SOURCES = ["capi.hip", "noalloc.hip", "plan.hip", "tables.hip", "field.hip", "farfield.hip", "assemble.hip",
           "generic.hip", "lu.hip"]
HOST_SOURCES = ["rwg.cpp"]   # host-only C++ (no device code): g++
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BIELCC_ARCH", "gfx950")
# packed FP32 instructions are disabled for all device code
CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def _compile_host(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"), "-c",
           os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stderr}")
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES)) + list(ex.map(_compile_host, HOST_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lrocblas",
                                                                                    "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_example()
    if verbose:
        print(f"built {LIB}")
    return LIB


EXAMPLE = os.path.join(HERE, "lib", "dropin_example")


def build_cpp_example() -> str:
    """The drop-in C++ layer's example program (host g++ against the C ABI)."""
    src = os.path.join(HERE, "cpp", "examples", "dropin_example.cpp")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "cpp"),
           src, "-o", EXAMPLE, "-L", os.path.dirname(LIB), "-lbielcc_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for the C++ example:\n{r.stderr}")
    return EXAMPLE


if __name__ == "__main__":
    build()
    sys.exit(0)
