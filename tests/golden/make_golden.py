"""Writes the golden fixtures of this directory (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no test vectors for its solver, and Taichi cannot run
here, so the fixtures freeze what the oracle (the C restatement of mgpcg.py,
itself pinned by closed forms and brute force in tests/test_oracle_core.py)
computes on two small seeded systems:

  c1_2d_64   2D 64^2 pool + disc pressure system, 4 levels (the 2D scene at
             half size), ghost fluid off
  c2_3d_32   3D 32^3 pool + sphere ghost-fluid pressure system, 3 levels

Each file holds the inputs the solver receives (cell types, Adiag and Ax of
every level in the canonical layout, and b) and the outputs: x and the
iteration count of a PLAIN (FP32, the reference's arithmetic) and of a MIXED
solve at rel_tol 1e-12 (1e-6 relative residual), nu = 2, bottom = 10.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "taichi-ferrofluid_amd"))
sys.path.insert(0, ROOT)

CASES = {
    "c1_2d_64": lambda scenes: scenes.c1_2d(seed=0, n=64),
    "c2_3d_32": lambda scenes: scenes.c2_3d(seed=1, n=32, L=3),
}


def build(name):
    from mgpcg_amd import scenes
    from oracle import core
    from oracle import systems as osys

    sysd = CASES[name](scenes)
    h, b = osys.assemble(sysd)
    out = {"b": np.ascontiguousarray(b, np.float32), "L": np.int32(len(h["gt"]))}
    for l in range(len(h["gt"])):
        out[f"gt{l}"] = h["gt"][l]
        out[f"Adiag{l}"] = h["Adiag"][l]
        out[f"Ax{l}"] = h["Ax"][l]
    for prec, tag in ((core.PLAIN, "plain"), (core.MIXED, "mixed")):
        x, res, _ = core.solve(h, b, precision=prec, nu=2, bottom=10, max_iters=500,
                               rel_tol=1e-12)
        out[f"x_{tag}"] = x if prec == core.MIXED else x.astype(np.float32)
        out[f"iters_{tag}"] = np.int32(res["iters"])
        out[f"status_{tag}"] = np.int32(res["status"])
    return out


def main():
    for name in CASES:
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **build(name))
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
