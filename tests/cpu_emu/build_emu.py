"""Build the CPU emulation of libmgpcg -- TEST INFRASTRUCTURE ONLY.

Compiles the library's own HIP sources (taichi-ferrofluid_amd/csrc) as host
C++ against the emulated runtime in this directory, into libmgpcg_emu.so,
which the CPU tests load from Python (guard-page allocator) to run every
kernel and the engine host code.  The product never loads it.

build_if_stale() rebuilds when the sources changed since the last build (a
hash of every source, header and this recipe is kept beside the library).
"""
import fcntl
import glob
import hashlib
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
STAMP = LIB + ".sha256"

CXXFLAGS = ["-std=c++17", "-O1", "-g", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-Wno-unused-result", "-Wno-unused-function", "-Wno-attributes",
            "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _inputs():
    files = sorted(glob.glob(os.path.join(CSRC, "*")) +
                   glob.glob(os.path.join(HERE, "include", "hip", "*.h")) +
                   glob.glob(os.path.join(ROOT, "include", "*.h")) +
                   [RUNTIME, os.path.abspath(__file__)])
    return [f for f in files if os.path.isfile(f)]


def source_hash():
    h = hashlib.sha256()
    for f in _inputs():
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def build(verbose=False):
    digest = source_hash()
    cmd = ["g++", *CXXFLAGS, "-shared", "-x", "c++", *SOURCES, "-x", "none", RUNTIME,
           "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    with open(STAMP, "w") as fh:
        fh.write(digest + "\n")
    return LIB


def build_if_stale(verbose=False):
    """Build unless the library on disk was built from the current sources
    (under a file lock: parallel test workers build it once)."""
    with open(LIB + ".lock", "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        if os.path.exists(LIB) and os.path.exists(STAMP):
            with open(STAMP) as fh:
                if fh.read().strip() == source_hash():
                    return LIB
        return build(verbose=verbose)


if __name__ == "__main__":
    print(build(verbose=True))
