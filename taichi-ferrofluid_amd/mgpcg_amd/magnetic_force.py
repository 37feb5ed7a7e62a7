"""MagneticForceStrategy (magnetic_force.py:6-39).

The pressure strategy whose AIR faces carry the magnetic pressure jump
p0 + 1/2 k mu0 |H|^2, with H interpolated at the face (H_interp,
ferrofluid_simulator.py:66-70) and no ghost-fluid weight in b; A is inherited.
``H`` is a list of dim float32 face fields (shape res + e_k).
"""
from . import _lib
from .pressure_project import PressureProjectStrategy


class MagneticForceStrategy(PressureProjectStrategy):
    def __init__(self, dim, velocity, ghost_fluid_method, phi, p0, mu0, k, H, simulator=None,
                 c_max=1e3):
        super().__init__(dim, velocity, ghost_fluid_method, phi, p0, c_max=c_max)
        self.mu0 = mu0
        self.k = k
        self.H = H
        self.simulator = simulator

    def build_b(self, solver):
        """build_b_kernel (magnetic_force.py:17-39)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_magnetic_build_b(
            _lib.stream_handle(solver.device), self.dim, _lib.int3(solver.res),
            _lib.ptr(solver.grid_type[0]), _lib.ptr_array(self.velocity), _lib.ptr_array(self.H),
            float(self.scale_A), float(self.scale_b), float(self.p0), float(self.mu0),
            float(self.k), _lib.ptr(solver.b)))
