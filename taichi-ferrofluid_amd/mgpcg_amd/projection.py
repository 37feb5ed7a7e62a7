"""The step after the pressure solve: the gradient applied to the MAC velocity,
and the velocity extrapolation that follows it.

apply_pressure follows FluidSimulator.apply_pressure (fluid_simulator.py:192-214),
apply_magnetic_force FerrofluidSimulator.apply_magnetic_force
(ferrofluid_simulator.py:114-137) and extrap_velocity
FluidSimulator.extrap_velocity (fluid_simulator.py:240-270).  All run in the
HIP library and update ``velocity`` (a list of dim float32 device tensors,
component k of shape res + e_k) in place; ``pressure`` is the solver's x.
"""
import torch

from . import _lib


def _vel_ptrs(velocity):
    for u in velocity:
        assert u.is_contiguous() and u.dtype.is_floating_point
    return _lib.ptr_array(velocity)


def apply_pressure(velocity, cell_type, pressure, dt, rho, dx, p0, ghost_fluid_method=False,
                   phi=None, c_max=1e3):
    """fluid_simulator.py:192-214.  ``cell_type`` is int8 (FLUID 0, AIR 1,
    SOLID 2); ``phi`` is needed with the ghost-fluid method."""
    lib = _lib.load()
    dim = len(velocity)
    res = list(cell_type.shape)
    dev = cell_type.device
    _lib.check(lib.mgpcg_apply_pressure(
        _lib.stream_handle(dev), dim, _lib.int3(res), _lib.ptr(cell_type),
        _lib.ptr(phi) if phi is not None else None, _lib.ptr(pressure), _vel_ptrs(velocity),
        float(dt), float(rho), float(dx), float(p0), int(bool(ghost_fluid_method)),
        float(c_max)))


def apply_magnetic_force(velocity, cell_type, pressure, H, dt, rho, dx, p0, mu0, k):
    """ferrofluid_simulator.py:114-137.  ``H`` is the face-centred field
    (component a of shape res + e_a)."""
    lib = _lib.load()
    dim = len(velocity)
    res = list(cell_type.shape)
    dev = cell_type.device
    _lib.check(lib.mgpcg_apply_magnetic_force(
        _lib.stream_handle(dev), dim, _lib.int3(res), _lib.ptr(cell_type), _lib.ptr(pressure),
        _lib.ptr_array(H), _vel_ptrs(velocity), float(dt), float(rho), float(dx), float(p0),
        float(mu0), float(k)))


def extrap_velocity(velocity, cell_type, iterations=10):
    """fluid_simulator.py:264-270: extend the velocity from the faces next to
    FLUID cells outwards by `iterations` layers."""
    lib = _lib.load()
    dim = len(velocity)
    res = list(cell_type.shape)
    dev = cell_type.device
    n = max(u.numel() for u in velocity)
    backup = torch.empty(n, dtype=torch.float32, device=dev)
    valid = torch.empty(n, dtype=torch.int8, device=dev)
    valid_temp = torch.empty(n, dtype=torch.int8, device=dev)
    _lib.check(lib.mgpcg_extrap_velocity(
        _lib.stream_handle(dev), dim, _lib.int3(res), _lib.ptr(cell_type), _vel_ptrs(velocity),
        int(iterations), _lib.ptr(backup), _lib.ptr(valid), _lib.ptr(valid_temp)))

