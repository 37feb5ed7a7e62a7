"""SurfaceTension and SurfaceTensionStrategy (surface_tension.py:10-153).

A semi-implicit capillary solve per velocity axis d: A = I/dt plus a
sigma dt / dx_l^2 delta(phi)-weighted Laplacian over the whole box, b from the
velocity, the normal n and the curvature kappa.  The strategy reads the axis
from ``self.d`` at call time, so each axis gets its own system (the reference
sets ``self.d`` per axis, surface_tension.py:142).

``simulator`` needs: dim, res, dt, dx, real, velocity (list of MAC tensors),
level_set (with .phi), poisson_solver, verbose.
"""
import time

import torch

from . import _lib
from . import utils


class SurfaceTensionStrategy:
    writes_all_cells = True

    def __init__(self, simulator, surface_tension):
        self.dim = simulator.dim
        self.res = list(simulator.res)
        self.dt = simulator.dt
        self.dx = simulator.dx
        self.real = simulator.real
        self.sigma = surface_tension.sigma
        self.simulator = simulator
        self.surface_tension = surface_tension
        self.d = 0

    def build_b(self, solver):
        """build_b_kernel (surface_tension.py:23-66)."""
        lib = _lib.load()
        st = self.surface_tension
        _lib.check(lib.mgpcg_surface_build_b(
            _lib.stream_handle(solver.device), self.dim, _lib.int3(self.res), int(self.d),
            _lib.ptr(self.simulator.velocity[self.d]), _lib.ptr(self.simulator.level_set.phi),
            _lib.ptr(st.kappa), _lib.ptr(st.n), float(self.dt), float(self.dx),
            float(self.sigma), _lib.ptr(solver.b)))

    def build_A(self, solver, level):
        """build_A_kernel (surface_tension.py:71-89)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_surface_build_A(
            _lib.stream_handle(solver.device), self.dim, _lib.int3(self.res), int(level),
            int(self.d), _lib.ptr(self.simulator.level_set.phi), float(self.dt), float(self.dx),
            float(self.sigma), _lib.ptr(solver.Adiag[level]), _lib.ptr(solver.Ax[level])))


class SurfaceTension:
    def __init__(self, simulator, sigma=7.28e-2, iterations=500):
        self.dim = simulator.dim
        self.res = list(simulator.res)
        self.dx = simulator.dx
        self.real = simulator.real
        self.sigma = sigma  # surface tension coefficient (surface_tension.py:102)
        self.simulator = simulator
        self.level_set = simulator.level_set
        self.poisson_solve_iterations = iterations
        self.poisson_solver = simulator.poisson_solver
        dev = self.level_set.phi.device
        self.n = torch.zeros(self.res + [self.dim], dtype=torch.float32, device=dev)
        self.kappa = torch.zeros(self.res, dtype=torch.float32, device=dev)
        self.strategy = SurfaceTensionStrategy(simulator, self)

    def calc_n(self, d, phi):
        """calc_n (surface_tension.py:115-127)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_calc_n(_lib.stream_handle(phi.device), self.dim,
                                    _lib.int3(self.res), int(d), _lib.ptr(phi), float(self.dx),
                                    _lib.ptr(self.n)))

    def calc_kappa(self, d, phi):
        """calc_kappa (surface_tension.py:129-136)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_calc_kappa(_lib.stream_handle(phi.device), self.dim,
                                        _lib.int3(self.res), _lib.ptr(self.n), float(self.dx),
                                        _lib.ptr(self.kappa)))

    def solve_surface_tension(self):
        """surface_tension.py:138-153."""
        verbose = getattr(self.simulator, "verbose", False)
        for k in range(self.dim):
            self.calc_n(k, self.level_set.phi)
            self.calc_kappa(k, self.level_set.phi)
            self.strategy.d = k

            start1 = time.perf_counter()
            self.poisson_solver.full_reinitialize(self.strategy)
            end1 = time.perf_counter()

            start2 = time.perf_counter()
            self.poisson_solver.solve(self.poisson_solve_iterations, verbose)
            end2 = time.perf_counter()

            if verbose:
                print(f"surface tension, axis {k}: assembly {1e3 * (end1 - start1):.3f} ms, "
                      f"solve {1e3 * (end2 - start2):.3f} ms")
            utils.copy(self.poisson_solver.x, self.simulator.velocity[k][
                tuple(slice(0, n) for n in self.res)])
