"""Builds libcassandra_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m cassandra_amd.build          # or cassandra_amd.build.build()
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcassandra_amd.so")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["patterns.cpp", "runtime.hip", "solver.hip", "evaluator.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CAD_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # loops marked `#pragma unroll` unroll fully also when large (the
    # matrix-free gather's slot loop indexes register arrays; partly unrolled,
    # they would live in scratch memory)
    "-mllvm", "-pragma-unroll-threshold=100000",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


CPP_DIR = os.path.join(HERE, "cpp")
CPP_SOURCES = ["device.cpp", "Vector.cpp", "MatrixCSR.cpp", "Board.cpp", "Helpers.cpp", "Indexer.cpp",
               "Evaluator.cpp", "PatternFit.cpp", "Binary.cpp"]
CPP_OUT = os.path.join(HERE, "libcassandra_fit.so")
TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "dropin_test.cpp")
TEST_OUT = os.path.join(OBJ, "dropin_test")
GXX = os.environ.get("CXX", "g++")
CPP_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-fopenmp", "-Wall", f"-I{os.path.join(CPP_DIR, 'include')}"]


def build_cpp(verbose: bool = False) -> str:
    """The reference-API C++ layer (libcassandra_fit.so) over the C-ABI, and
    its test runner."""
    os.makedirs(OBJ, exist_ok=True)

    def cc(src):
        obj = os.path.join(OBJ, "cpp_" + os.path.splitext(src)[0] + ".o")
        r = subprocess.run([GXX, *CPP_FLAGS, "-c", os.path.join(CPP_DIR, "src", src), "-o", obj],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stderr[-6000:]}")
        return obj

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(cc, CPP_SOURCES))
    link = ["-L" + HERE, "-lcassandra_amd", "-Wl,-rpath,$ORIGIN", "-fopenmp"]
    r = subprocess.run([GXX, "-shared", *objs, "-o", CPP_OUT + ".tmp", *link], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(CPP_OUT + ".tmp", CPP_OUT)
    r = subprocess.run([GXX, *CPP_FLAGS, TEST_SRC, "-o", TEST_OUT, "-L" + HERE, "-lcassandra_fit", "-lcassandra_amd",
                        f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"test runner build failed:\n{r.stderr[-6000:]}")
    if verbose:
        print("built", CPP_OUT, "and", TEST_OUT)
    return CPP_OUT


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_cpp(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
