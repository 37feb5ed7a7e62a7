"""Turn a scene dict (scenes.py) into an assembled system on the device.

This is the per-request front end shared by the solver wrapper, the batch
engine and the server: upload the scene's fields, build the matching
strategy and run ``reinitialize`` / ``full_reinitialize`` (mgpcg.py:79-94).
"""
import types

import numpy as np
import torch

from . import _lib
from .level_set import LevelSet
from .potential_function import PotentialFunctionStrategy
from .pressure_project import PressureProjectStrategy
from .surface_tension import SurfaceTension
from .volume_control import PressureProjectWithVolumeControlStrategy


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


class SeededVolumeControl(PressureProjectWithVolumeControlStrategy):
    """Volume-control pressure system with a given divergence constant c
    (the value the PI controller of volume_control.py:45-50 would produce)."""

    def __init__(self, *args, c=0.0, **kw):
        super().__init__(*args, **kw)
        self.c = c

    def build_b(self, solver):
        PressureProjectStrategy.build_b(self, solver)
        self.add_c(solver, self.c)

    def packed_c(self):
        return self.c


def upload(sysd, device):
    """Device copies of every array field of a scene."""
    out = {}
    for k, v in sysd.items():
        if isinstance(v, np.ndarray):
            out[k] = _t(v, device)
        elif isinstance(v, list) and v and isinstance(v[0], np.ndarray):
            out[k] = [_t(a, device) for a in v]
        else:
            out[k] = v
    return out


def assemble(solver, dev, packed=False):
    """Build b and the operator hierarchy of scene ``dev`` (uploaded) in ``solver``.

    ``packed`` assembles the pressure kinds straight into the device hierarchy
    (reinitialize_packed); the other kinds take the canonical route either way.
    """
    kind = dev["kind"]
    dim, res = dev["dim"], dev["res"]
    if kind in ("pressure", "volume"):
        if kind == "pressure":
            st = PressureProjectStrategy(dim, dev["velocity"], dev["gfm"], dev["phi"], dev["p0"],
                                         c_max=dev["c_max"])
        else:
            ls = LevelSet(dim, res, dev["dx"], phi=dev["phi"], device=solver.device)
            st = SeededVolumeControl(dim, dev["velocity"], dev["gfm"], dev["phi"], dev["p0"], ls,
                                     dev["dt"], c_max=dev["c_max"], c=dev["c"])
        st.scale_A = dev["scale_A"]
        st.scale_b = dev["scale_b"]
        if packed:
            solver.reinitialize_packed(dev["cell_type"], st)
        else:
            solver.reinitialize(dev["cell_type"], st)
        return st
    if kind == "psi":
        st = PotentialFunctionStrategy(dim, res, dev["dx"], dev["H_ext"], dev["chi"])
        solver.full_reinitialize(st)
        return st
    if kind == "surface":
        sim = types.SimpleNamespace(
            dim=dim, res=res, dt=dev["dt"], dx=dev["dx"], real=float, velocity=dev["velocity"],
            level_set=LevelSet(dim, res, dev["dx"], phi=dev["phi"], device=solver.device),
            poisson_solver=solver, verbose=False)
        st_obj = SurfaceTension(sim, sigma=dev.get("sigma", 7.28e-2))
        d = dev["d"]
        st_obj.calc_n(d, dev["phi"])
        st_obj.calc_kappa(d, dev["phi"])
        st_obj.strategy.d = d
        solver.full_reinitialize(st_obj.strategy)
        return st_obj
    raise ValueError(f"unknown system kind {kind!r}")


def solve_params(sysd):
    """(max_iters, rel_tol) of the reference call site of this system kind."""
    # every call site uses 500 iterations (fluid_simulator.py:95, surface_tension.py:107);
    # the measured workloads use rel_tol 1e-12 on squared norms (1e-6 relative residual)
    return 500, 1e-12


__all__ = ["assemble", "upload", "solve_params", "SeededVolumeControl", "_lib"]
