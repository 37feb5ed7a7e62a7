"""The parts of LevelSet (level_set.py:10-37) the solver's strategies use.

``theta`` and ``delta`` (level_set.py:24-37) are evaluated inside the HIP
assembly kernels; this class carries the level set ``phi`` and its ``dx``
(the interface thickness is eps = 5 dx, level_set.py:22).  Redistancing
(fast marching / fast sweeping, level_set.py:166-392) is a separate workload
and not part of the solver path.
"""
import math

import torch


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
