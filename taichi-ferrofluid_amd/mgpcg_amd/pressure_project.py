"""PressureProjectStrategy (pressure_project.py:5-70) over the HIP assembly kernels.

Same constructor and protocol as the reference: the caller sets ``scale_A``
(dt / (rho dx^2)) and ``scale_b`` (1 / dx) (fluid_simulator.py:178-179), then
``solver.reinitialize(cell_type, strategy)`` calls ``build_b(solver)`` and
``build_A(solver, level)``.  ``velocity`` is a list of dim float32 MAC
tensors (shape res + e_k), ``phi`` the float32 level set [*res].

``c_max`` is the ghost-fluid clamp ``min(c, 1e3)`` of pressure_project.py:37,62
(1e3 in the reference; a clamp theta >= theta_min is c_max = 1 / theta_min).

``assemble_packed`` is the serving path's route to the same system: b and
every level's operator computed straight into the solver's device hierarchy
(mgpcg_assemble_pressure), bit for bit what reinitialize + load produce,
without writing the canonical ``b``, ``Adiag`` and ``Ax`` fields.
"""
from . import _lib


class PressureProjectStrategy:
    writes_all_cells = True  # the kernels write every cell of b, Adiag, Ax

    def __init__(self, dim, velocity, ghost_fluid_method, phi, p0, c_max=1e3):
        self.dim = dim
        self.velocity = velocity
        self.ghost_fluid_method = ghost_fluid_method
        self.phi = phi
        self.p0 = p0  # the standard atmospheric pressure
        self.c_max = c_max
        self.scale_A = 1.0
        self.scale_b = 1.0

    def build_b(self, solver):
        """build_b_kernel (pressure_project.py:15-39)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_pressure_build_b(
            _lib.stream_handle(solver.device), self.dim, _lib.int3(solver.res),
            _lib.ptr(solver.grid_type[0]), _lib.ptr_array(self.velocity), _lib.ptr(self.phi),
            float(self.scale_A), float(self.scale_b), float(self.p0),
            int(bool(self.ghost_fluid_method)), float(self.c_max), _lib.ptr(solver.b)))

    def packed_c(self):
        """The divergence constant added to b on FLUID cells, or None (none
        for the plain pressure system)."""
        return None

    def assemble_packed(self, solver):
        """b and every level's operator straight into the device hierarchy
        (solver.grid_type[0] holds the cell types; coarser levels' are
        written here)."""
        c = self.packed_c()
        lib = _lib.load()
        _lib.check(lib.mgpcg_set_stream(solver._h, _lib.stream_handle(solver.device)))
        _lib.check(lib.mgpcg_assemble_pressure(
            solver._h, _lib.ptr_array(solver.grid_type), _lib.ptr_array(self.velocity),
            _lib.ptr(self.phi) if self.phi is not None else None, float(self.scale_A),
            float(self.scale_b), float(self.p0), int(bool(self.ghost_fluid_method)),
            float(self.c_max), int(c is not None), float(c or 0.0)))

    def build_A(self, solver, level):
        """build_A_kernel (pressure_project.py:45-64)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_pressure_build_A(
            _lib.stream_handle(solver.device), self.dim, _lib.int3(solver.res), int(level),
            _lib.ptr(solver.grid_type[level]), _lib.ptr(self.phi), float(self.scale_A),
            int(bool(self.ghost_fluid_method)), float(self.c_max), _lib.ptr(solver.Adiag[level]),
            _lib.ptr(solver.Ax[level])))
