"""Build recipe for the C oracle (test infrastructure only).

Compiles oracle/mgpcg_oracle.c into oracle/liboracle.so with gcc.
``-ffp-contract=off`` keeps every multiply and add separately rounded,
matching the HIP kernels' operation order; OpenMP parallelises the
order-independent sweeps (red-black half-sweeps, gather restriction) and the
chunked FP64 dot, whose result does not depend on the thread count.

The reference (Python/Taichi) has no compilable C/C++ sources, so there is no
``oracle/_ref`` build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "mgpcg_oracle.c")
OUT = os.path.join(HERE, "liboracle.so")

CFLAGS = [
    "-O3",
    "-march=x86-64-v3",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fopenmp",
    "-fPIC",
    "-shared",
    "-std=c11",
    "-Wall",
    "-Wno-unknown-pragmas",
]


def build(verbose=False):
    cmd = ["gcc", *CFLAGS, SRC, "-o", OUT + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
