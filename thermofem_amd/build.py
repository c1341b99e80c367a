"""Builds the gfx950 HIP library thermofem_amd/libthermofem_hip.so in-tree.

Every .hip source under thermofem_amd/csrc is compiled with hipcc for gfx950
(in parallel), then linked into one shared library whose C ABI is declared
in include/thermofem_hip.h.  Packed-FP32 instruction selection is disabled
(-packed-fp32-ops) for all device code.  The host programs over the C++20
headers (the solve-matrix CLI and the C++ API tests) are then compiled with
g++ -std=c++20 and linked against the library.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libthermofem_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TF_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-fast-math", "-I" + os.path.join(HERE, "..", "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "bin")
CXX = os.environ.get("CXX", "g++")
# Host programs over the C++20 headers, linked against the library.
PROGRAMS = {
    os.path.join(BIN, "solve-matrix"): os.path.join(HERE, "cli", "solve_matrix.cpp"),
    os.path.join(ROOT, "tests", "cpp", "build", "api_test"): os.path.join(ROOT, "tests", "cpp", "api_test.cpp"),
}


def _program(out, src):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    rpath = os.path.relpath(HERE, os.path.dirname(out))
    cmd = [CXX, "-std=c++20", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(HERE, "include"), "-I" + os.path.dirname(src), src, "-o", out + ".tmp",
           "-L" + HERE, "-lthermofem_hip", f"-Wl,-rpath,$ORIGIN/{rpath}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{CXX} failed for {src}:\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build_programs(verbose=False):
    with cf.ThreadPoolExecutor(len(PROGRAMS)) as ex:
        outs = list(ex.map(lambda kv: _program(*kv), PROGRAMS.items()))
    if verbose:
        for o in outs:
            print("built", o)
    return outs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_programs(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
