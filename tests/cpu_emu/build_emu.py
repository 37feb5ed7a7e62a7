"""Build the CPU emulation of libmgpcg -- TEST INFRASTRUCTURE ONLY.

Compiles the library's own HIP sources (taichi-ferrofluid_amd/csrc) as host
C++ against the emulated runtime in this directory:
  libmgpcg_emu.so   loadable from Python (guard-page allocator), used by the
                    CPU tests to run every kernel and the engine host code;
  emu_driver        (optional, --asan) an AddressSanitizer build of a C++
                    driver over the same C-ABI.
The product never loads these.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "taichi-ferrofluid_amd", "csrc")
SOURCES = [os.path.join(CSRC, s) for s in
           ("mgpcg_solver.hip", "mgpcg_strategies.hip", "mgpcg_levelset.hip")]
RUNTIME = os.path.join(HERE, "emu_runtime.cpp")
LIB = os.path.join(HERE, "libmgpcg_emu.so")

CXXFLAGS = ["-std=c++17", "-O1", "-g", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-Wno-unused-result", "-Wno-unused-function", "-Wno-attributes",
            "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def build(verbose=False):
    cmd = ["g++", *CXXFLAGS, "-shared", "-x", "c++", *SOURCES, "-x", "none", RUNTIME,
           "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_asan_driver(verbose=False):
    out = os.path.join(HERE, "emu_driver")
    cmd = ["g++", *CXXFLAGS, "-fsanitize=address", "-fno-omit-frame-pointer", "-x", "c++",
           *SOURCES, "-x", "none", RUNTIME, os.path.join(HERE, "emu_driver.cpp"),
           os.path.join(ROOT, "oracle", "mgpcg_oracle.c"), "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    if "--asan" in sys.argv:
        print(build_asan_driver(verbose=True))
