"""LevelSet (level_set.py:10-37) and FastSweepingLevelSet.redistance
(level_set.py:260-333).

``theta`` and ``delta`` (level_set.py:24-37) are evaluated inside the HIP
assembly kernels; this class carries the level set ``phi`` and its ``dx``
(the interface thickness is eps = 5 dx, level_set.py:22).  Redistancing runs
in the HIP library (target_surface, then repeat_times x six directional
sweeps of the Eikonal update; neighbouring lines are read as they were
before each sweep, which makes the result deterministic).
"""
import math

import torch

from . import _lib


class LevelSet:
    def __init__(self, dim, res, dx, real=float, phi=None, device=None):
        self.dim = dim
        self.res = list(res)
        self.dx = dx
        self.real = real
        dev = torch.device(device) if device is not None else torch.device("cuda")
        self.phi = phi if phi is not None else torch.zeros(self.res, dtype=torch.float32,
                                                           device=dev)
        self.eps = 5 * self.dx

    def theta(self, phi):
        """Smoothed Heaviside (level_set.py:24-30), elementwise on a tensor."""
        eps = self.eps
        mid = (0.5 + phi / (2 * eps)) + (1 / (2 * math.pi)) * torch.sin(math.pi * phi / eps)
        return torch.where(phi <= -eps, torch.zeros_like(phi),
                           torch.where(phi >= eps, torch.ones_like(phi), mid))

    def delta(self, phi):
        """Smoothed Dirac delta (level_set.py:33-37), elementwise on a tensor."""
        eps = self.eps
        mid = (1 + torch.cos(math.pi * phi / eps)) / (2 * eps)
        return torch.where((phi <= -eps) | (phi >= eps), torch.zeros_like(phi), mid)


class FastSweepingLevelSet(LevelSet):
    """level_set.py:260-333 (repeat_times = 2, level_set.py:264)."""

    def __init__(self, dim, res, dx, real=float, phi=None, device=None):
        super().__init__(dim, res, dx, real, phi, device)
        self.repeat_times = 2
        dev = self.phi.device
        self.phi_temp = torch.empty_like(self.phi)
        self._snap = torch.empty_like(self.phi)
        self._fixed = torch.empty(self.res, dtype=torch.int8, device=dev)

    def redistance(self):
        """level_set.py:329-333: phi <- signed distance to its zero crossing."""
        lib = _lib.load()
        _lib.check(lib.mgpcg_redistance(
            _lib.stream_handle(self.phi.device), self.dim, _lib.int3(self.res), float(self.dx),
            int(self.repeat_times), _lib.ptr(self.phi), _lib.ptr(self.phi_temp),
            _lib.ptr(self._snap), _lib.ptr(self._fixed)))

