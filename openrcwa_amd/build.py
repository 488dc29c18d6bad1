"""Build the native engine: every HIP/C++ source under ``csrc/`` is compiled
for gfx950 with hipcc and linked into ``libopenrcwa_gpu.so`` next to this file.

Usage:  python -m openrcwa_amd.build [-v]
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libopenrcwa_gpu.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for all device code
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unknown-warning-option"]

# .hip files hold device code; .cpp files are host-only C++ (no HIP headers)
SOURCES = ["blas.hip", "eig.hip", "rcwa_kernels.hip", "field_kernels.hip", "plan.hip",
           "smatrix.hip", "mlp.hip", "debug.hip", "ofd_bridge.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    if verbose and r.stdout.strip():
        print(r.stdout)
    return r


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile every source (always, no staleness check) and link the library."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD_DIR, os.path.splitext(s)[0] + ".o") for s in srcs]

    def compile_one(i):
        lang = "c++" if srcs[i].endswith(".cpp") else "hip"
        _run([HIPCC, *FLAGS, "-x", lang, "-c", os.path.join(CSRC, srcs[i]), "-o", objs[i]],
             verbose)

    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        list(ex.map(compile_one, range(len(srcs))))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    build_cpp_harness(verbose)
    return LIB


ROOT = os.path.dirname(HERE)
HARNESS_SRC = os.path.join(ROOT, "tests", "cpp_facade_harness.cpp")
HARNESS_BIN = os.path.join(ROOT, "tests", "_bin", "cpp_facade_harness")


def build_cpp_harness(verbose: bool = False) -> str:
    """A C++ caller of include/openrcwa_gpu.hpp linked against the library
    (tests/cpp_facade_harness.cpp; the GPU tests run it)."""
    if not os.path.exists(HARNESS_SRC):
        return ""
    os.makedirs(os.path.dirname(HARNESS_BIN), exist_ok=True)
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), HARNESS_SRC,
          "-L" + HERE, "-lopenrcwa_gpu", "-Wl,-rpath,$ORIGIN/../../openrcwa_amd",
          "-o", HARNESS_BIN], verbose)
    return HARNESS_BIN


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
