"""Assemble a scene dict (taichi-ferrofluid_amd/scenes.py) with the numpy
strategies -- TEST INFRASTRUCTURE ONLY.

Returns the hierarchy {gt, Adiag, Ax} and b in canonical layout, the CPU
counterpart of systems.assemble on the device.
"""
import numpy as np

from . import core
from . import strategies as S


def assemble(sysd):
    kind = sysd["kind"]
    dim, res, L = sysd["dim"], list(sysd["res"]), sysd["L"]
    if kind in ("pressure", "volume"):
        ct = np.ascontiguousarray(sysd["cell_type"], np.int8)
        gfm, c_max = sysd["gfm"], sysd["c_max"]
        phi = sysd["phi"]
        b = S.pressure_build_b(ct, sysd["velocity"], sysd["scale_A"], sysd["scale_b"], sysd["p0"],
                               gfm, phi, c_max)
        if kind == "volume":
            b = S.add_c(ct, b, sysd["c"])
        h = S.hierarchy_from_cell_type(
            ct, L, lambda gt, l: S.pressure_build_A(gt, l, sysd["scale_A"], gfm, phi, c_max))
        return h, b
    if kind == "psi":
        b = S.potential_build_b(res, sysd["chi"], sysd["H_ext"])
        h = S.hierarchy_full(res, L, lambda l: S.potential_build_A(res, l, sysd["dx"], sysd["chi"]))
        return h, b
    if kind == "surface":
        d = sysd["d"]
        n = S.calc_n(sysd["phi"], d, sysd["dx"])
        kappa = S.calc_kappa(n, sysd["dx"])
        b = S.surface_build_b(sysd["velocity"][d], sysd["phi"], kappa, n, d, sysd["dt"],
                              sysd["dx"], sysd.get("sigma", 7.28e-2))
        h = S.hierarchy_full(res, L, lambda l: S.surface_build_A(res, l, d, sysd["phi"],
                                                                  sysd["dt"], sysd["dx"],
                                                                  sysd.get("sigma", 7.28e-2)))
        return h, b
    raise ValueError(kind)


def solve(sysd, precision=core.PLAIN, **kw):
    h, b = assemble(sysd)
    if precision == core.F64:
        h = core.to_f64(h)
    kw.setdefault("nu", sysd.get("nu", 2))
    kw.setdefault("bottom", sysd.get("bottom", 10))
    return core.solve(h, b, precision=precision, **kw)
