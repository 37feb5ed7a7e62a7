"""Build recipe of libmgpcg.so (HIP, gfx950), in-tree.

The library is linked against the HIP runtime by soname (libamdhip64.so.7);
loaded after ``import torch`` it binds to the runtime torch already mapped,
so there is one HIP runtime per process.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)  # taichi-ferrofluid_amd/
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(os.path.dirname(PKG_ROOT), "include")
OUT = os.path.join(PKG_ROOT, "libmgpcg.so")

SOURCES = ["mgpcg_solver.hip", "mgpcg_strategies.hip", "mgpcg_levelset.hip"]

HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every multiply and add rounded separately: FP32 results reproduce a CPU
    # restatement with the same operation order bit for bit
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmd = ["hipcc", *HIPFLAGS, "-I", INCLUDE, "-I", CSRC, *srcs, "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError("hipcc failed building libmgpcg.so")
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
