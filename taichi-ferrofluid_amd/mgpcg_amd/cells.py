"""Cell classification before the solve, on the device.

apply_markers follows FluidSimulator.apply_markers (fluid_simulator.py:151-159):
every non-SOLID cell becomes FLUID where phi <= 0 and AIR elsewhere.
init_boundary follows FluidSimulator.init_boundary (fluid_simulator.py:337-341):
the one-cell border becomes SOLID.  ``cell_type`` is an int8 device tensor
(FLUID 0, AIR 1, SOLID 2), updated in place.
"""
from . import _lib


def apply_markers(cell_type, phi):
    lib = _lib.load()
    _lib.check(lib.mgpcg_apply_markers(_lib.stream_handle(cell_type.device), cell_type.dim(),
                                       _lib.int3(list(cell_type.shape)), _lib.ptr(phi),
                                       _lib.ptr(cell_type)))


def init_boundary(cell_type):
    lib = _lib.load()
    _lib.check(lib.mgpcg_init_boundary(_lib.stream_handle(cell_type.device), cell_type.dim(),
                                       _lib.int3(list(cell_type.shape)), _lib.ptr(cell_type)))
