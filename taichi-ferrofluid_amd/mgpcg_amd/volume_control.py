"""PressureProjectWithVolumeControlStrategy (volume_control.py:6-51).

The pressure strategy plus a PI-controlled divergence constant ``c`` added to
b on FLUID cells (Kim et al. 2007).  The controller state lives in Python as
in the reference; the volume integral runs on the device (FP64 accumulation
of the per-cell FP32 volumes, so the sum does not depend on launch order).
"""
import torch

from . import _lib
from .pressure_project import PressureProjectStrategy


class PressureProjectWithVolumeControlStrategy(PressureProjectStrategy):
    def __init__(self, dim, velocity, ghost_fluid_method, phi, p0, level_set, dt, c_max=1e3,
                 verbose=False):
        super().__init__(dim, velocity, ghost_fluid_method, phi, p0, c_max=c_max)
        self.level_set = level_set
        self.dt = dt
        self.verbose = verbose

        # Controller state (volume_control.py:13-20): the first call's volume
        # is the target; a PI controller with the reference's gains, k_p =
        # 2.3 / (n_p dt) with n_p = 25 steps, and k_i = (k_p / (2 zeta))^2
        # with damping ratio zeta = 2.
        self.step = 0
        self.vol_0 = 0.0
        self.y = 0.0  # time integral of the relative volume error
        self.n_p = 25
        self.k_p = 2.3 / (self.n_p * self.dt)
        self.zeta = 2
        self.k_i = (self.k_p / (2 * self.zeta)) ** 2
        self.last_c = None
        self._partials = None
        self._out = None

    def calc_volume(self):
        """calc_volume (volume_control.py:22-31) -> Python float."""
        phi = self.level_set.phi
        n = phi.numel()
        nb = (n + 255) // 256
        if self._partials is None or self._partials.numel() < nb:
            self._partials = torch.zeros(nb, dtype=torch.float64, device=phi.device)
            self._out = torch.zeros(1, dtype=torch.float64, device=phi.device)
        lib = _lib.load()
        _lib.check(lib.mgpcg_calc_volume(_lib.stream_handle(phi.device), self.dim,
                                         _lib.int3(list(phi.shape)), _lib.ptr(phi),
                                         float(self.level_set.dx), _lib.ptr(self._partials),
                                         int(self._partials.numel()), _lib.ptr(self._out)))
        return float(self._out.item())

    def add_c(self, solver, c):
        """add_c (volume_control.py:33-37)."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_add_c(_lib.stream_handle(solver.device), self.dim,
                                   _lib.int3(solver.res), _lib.ptr(solver.grid_type[0]), float(c),
                                   _lib.ptr(solver.b)))

    def _controller_c(self):
        """One controller step (volume_control.py:39-50): the constant to add,
        or None on the first call, which records the target volume."""
        self.step += 1
        vol = self.calc_volume()
        if self.step == 1:
            self.vol_0 = vol
            self.last_c = None
            return

        # relative error x, its integral y, and the divergence c the fluid
        # must gain per unit time to steer the volume back (volume_control.py:47-49)
        x = (vol - self.vol_0) / self.vol_0
        self.y += x * self.dt
        c = (-self.k_p * x - self.k_i * self.y) / (x + 1)
        self.last_c = c
        if self.verbose:
            print(f"volume control: step {self.step}, volume {vol:.6g} cells, c = {c:.6g}")
        return c

    def build_b(self, solver):
        """volume_control.py:39-51."""
        super().build_b(solver)
        c = self._controller_c()
        if c is not None:
            self.add_c(solver, c)

    def packed_c(self):
        return self._controller_c()
