"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py).

CPU: the scene generators, the numpy strategies and the oracle still produce
the fixtures bit for bit.  Device: the library, fed the fixture's inputs,
reproduces the PLAIN solution bit for bit with the same iteration count, and
the MIXED one within the solve tolerance; the device assembly of the same
scene reproduces the fixture's operator and right-hand side.
"""
import os
import sys

import numpy as np
import pytest
import torch

from helpers import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402

NAMES = sorted(make_golden.CASES)


def load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz")) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g = load(name)
    new = make_golden.build(name)
    assert sorted(new) == sorted(g)
    for k in g:
        assert new[k].dtype == g[k].dtype and np.array_equal(new[k], g[k]), k


def _solver_from_golden(g, dev, precision):
    from mgpcg_amd import mgpcg

    L = int(g["L"])
    res = list(g["gt0"].shape)
    s = mgpcg.MGPCGPoissonSolver(len(res), res, L, 2, 10, precision=precision, device=dev)
    for l in range(L):
        s.grid_type[l].copy_(torch.from_numpy(g[f"gt{l}"]))
        s.Adiag[l].copy_(torch.from_numpy(g[f"Adiag{l}"]))
        s.Ax[l].copy_(torch.from_numpy(g[f"Ax{l}"]))
    s.b.copy_(torch.from_numpy(g["b"]))
    return s


@pytest.mark.parametrize("name", NAMES)
def test_library_reproduces_golden_plain(dev, name):
    g = load(name)
    s = _solver_from_golden(g, dev, "plain")
    s.solve(500, rel_tol=1e-12)
    assert s.last["iters"] == int(g["iters_plain"])
    assert s.last["status"] == int(g["status_plain"])
    assert np.array_equal(s.x.cpu().numpy(), g["x_plain"])


@pytest.mark.parametrize("name", NAMES)
def test_library_reproduces_golden_mixed(dev, name):
    g = load(name)
    s = _solver_from_golden(g, dev, "mixed")
    s.solve(500, rel_tol=1e-12)
    assert s.last["iters"] == int(g["iters_mixed"])
    assert rel_l2(s.x64.cpu().numpy(), g["x_mixed"]) <= 1e-5


@pytest.mark.parametrize("name", NAMES)
def test_device_assembly_reproduces_golden(dev, name):
    from mgpcg_amd import mgpcg
    from mgpcg_amd import scenes
    from mgpcg_amd import systems

    g = load(name)
    sysd = make_golden.CASES[name](scenes)
    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], 2, 10, device=dev)
    systems.assemble(s, systems.upload(sysd, dev))
    assert np.array_equal(s.b.cpu().numpy(), g["b"])
    for l in range(int(g["L"])):
        assert np.array_equal(s.grid_type[l].cpu().numpy(), g[f"gt{l}"]), l
        assert np.array_equal(s.Adiag[l].cpu().numpy(), g[f"Adiag{l}"]), l
        assert np.array_equal(s.Ax[l].cpu().numpy(), g[f"Ax{l}"]), l
