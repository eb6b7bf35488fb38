"""Build the gfx950 group-by library in-tree.

Compiles ``cudf_amd/csrc/*.{hip,cpp}`` with hipcc for gfx950 and links
``cudf_amd/lib/libmig_groupby.so`` (the C-ABI library declared in
``include/mig_groupby.h``).  Run ``python -m cudf_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmig_groupby.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LIB = "/opt/rocm/lib"  # roctx (the timeline ranges of the host driver)
ARCH = "gfx950"

# -packed-fp32-ops: packed FP32 instructions are kept out of the generated code.
COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    f"-I{os.path.join(ROOT, 'include')}",
]

SOURCES = ["groupby_kernels.hip", "partition_kernels.hip", "partition_aggregate.hip", "batch_kernels.hip",
           "partitioning_kernels.hip", "groupby.cpp", "batch.cpp", "partitioning.cpp", "capi.cpp",
           "sort_kernels.hip", "sort_groupby.cpp", "groupby_noalloc.hip"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


CXX_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_cxx_api.cpp")
CXX_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_cxx_api")


def build_cxx_test() -> str:
    """The C++ API test: plain g++ against include/ (the cudf:: facade) and
    libmig_groupby.so; the HIP runtime only moves test data."""
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", f"-I{os.path.join(ROOT, 'include')}",
           "-I/opt/rocm/include", CXX_TEST_SRC, "-o", CXX_TEST_BIN, f"-L{LIBDIR}", "-lmig_groupby",
           "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN/../../cudf_amd/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ API test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return CXX_TEST_BIN


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, f"-L{ROCM_LIB}",
           "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM_LIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cxx_test()
    if verbose:
        print(f"built {LIB} and {CXX_TEST_BIN}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
