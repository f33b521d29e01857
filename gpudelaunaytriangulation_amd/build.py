"""Build libgpudelaunay.so (gfx950) in-tree with hipcc.

    python -m gpudelaunaytriangulation_amd.build

Every HIP source is compiled with FP contraction off (the FP64 path and the
FP64 fallback of the filtered path must round identically) and with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libgpudelaunay.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GDT_OFFLOAD_ARCH", "gfx950")

SOURCES = ["celllist.hip", "onering.hip", "testrepair.hip", "voronoi.hip", "mdstep.hip",
           "capi.hip", "noalloc.hip"]
# The LDS-window kernels, once per tile edge: (source, object, defines).
TILE_SOURCES = [(src, f"{src[:-4]}_ts{ts}", [f"-DGDT_TILE_SHIFT={ts}", f"-DGDT_WIN_MAXPTS={pts}"])
                for src in ("tile.hip", "ringtest.hip") for ts, pts in ((4, 1024), (5, 2048))]
# CPU mode: the same per-point routines compiled for the host (g++, OpenMP)
HOST_SOURCES = ["hostpath.cpp"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-command-line-argument",
    f"-I{CSRC}", f"-I{INCLUDE}",
]


HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
              f"-I{CSRC}", f"-I{INCLUDE}"]


def _compile(item) -> str:
    src, name, defines = item if isinstance(item, tuple) else (item, os.path.splitext(item)[0], [])
    obj = os.path.join(BUILD, name + ".o")
    if src.endswith(".cpp"):
        cmd = [os.environ.get("CXX", "g++"), *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


CXX = os.environ.get("CXX", "g++")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
BIN = os.path.join(PKG, "bin")
DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp")
DROPIN_BIN = os.path.join(BIN, "dropin_check")
GPUARRAY_SRC = os.path.join(ROOT, "tests", "cpp", "gpuarray_check.cpp")
GPUARRAY_BIN = os.path.join(BIN, "gpuarray_check")


def cxx_dropin_cmd(src: str, out: str) -> list[str]:
    """Host compile + link of a C++ caller of include/gpudelaunay.hpp."""
    return [CXX, "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
            f"-I{INCLUDE}", src, f"-L{PKG}", "-lgpudelaunay", f"-L{ROCM}/lib", "-lamdhip64",
            "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM}/lib", "-o", out]


def build_dropin_check(src: str = DROPIN_SRC, out: str = DROPIN_BIN) -> str:
    """A C++ drop-in check program (run by the GPU tests)."""
    os.makedirs(BIN, exist_ok=True)
    r = subprocess.run(cxx_dropin_cmd(src, out + ".tmp"), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ drop-in check {os.path.basename(src)} failed to build:\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library (and the C++
    drop-in check program against it); returns the library's path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES + TILE_SOURCES + HOST_SOURCES
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-lgomp", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    for src, out in ((DROPIN_SRC, DROPIN_BIN), (GPUARRAY_SRC, GPUARRAY_BIN)):
        if os.path.exists(src):
            build_dropin_check(src, out)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
