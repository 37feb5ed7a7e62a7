"""PotentialFunctionStrategy (potential_function.py:5-60).

The magnetic-potential system: a variable-coefficient (1 + chi) Laplacian
over the whole box (every level all FLUID, via full_reinitialize), pinned at
the reference point psi(0) = 0.  A is negative definite apart from the
reference row, so nothing in the solver assumes Adiag > 0 or pAp > 0.
``chi`` is float32 [*res]; ``H_ext`` a list of dim face fields (res + e_k).
"""
from . import _lib


class PotentialFunctionStrategy:
    writes_all_cells = True

    def __init__(self, dim, res, dx, H_ext, chi):
        self.dim = dim
        self.res = list(res)
        self.dx = dx
        self.H_ext = H_ext
        self.chi = chi

    def build_b(self, solver):
        """build_b_kernel (potential_function.py:19-34)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_potential_build_b(
            _lib.stream_handle(solver.device), self.dim, _lib.int3(self.res), _lib.ptr(self.chi),
            _lib.ptr_array(self.H_ext), _lib.ptr(solver.b)))

    def build_A(self, solver, level):
        """build_A_kernel (potential_function.py:40-55)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_potential_build_A(
            _lib.stream_handle(solver.device), self.dim, _lib.int3(self.res), int(level),
            float(self.dx), _lib.ptr(self.chi), _lib.ptr(solver.Adiag[level]),
            _lib.ptr(solver.Ax[level])))
