"""MGPCGPoissonSolver: the reference interface (mgpcg.py:4-238) over libmgpcg.

Public fields keep the reference's names and canonical layout, as torch
tensors on the GPU: ``b``, ``x`` (float32 [*res]), ``grid_type[l]`` (int8),
``Adiag[l]`` (float32) and ``Ax[l]`` (float32 [*res_l, dim]) with
``res_l = res // 2**l`` (mgpcg.py:20-40).  Strategies fill them exactly as the
reference strategies fill the Taichi fields (``build_b(solver)``,
``build_A(solver, level)``, pressure_project.py:41-70), so the strategy
modules of this package call the solver unchanged.

Everything numeric runs in the HIP library: ``solve`` loads the hierarchy
into its colour-split layout and runs the whole PCG loop on the device.
There is no CPU path.
"""
import ctypes
import time

import torch

from . import _lib
from . import utils

PRECISIONS = {"plain": _lib.PLAIN, "mixed": _lib.MIXED, "f64": _lib.F64}


def level_shape(res, l):
    return [n // (2 ** l) for n in res]


def pool_bytes(res, n_levels, precision="plain"):
    """Device bytes of the library's pool for one hierarchy of this shape,
    computed by the library from the layout it allocates (mgpcg_pool_bytes)."""
    n = ctypes.c_int64()
    _lib.check(_lib.load().mgpcg_pool_bytes(len(res), _lib.int3(list(res)), n_levels,
                                            PRECISIONS[precision], ctypes.byref(n)))
    return n.value


def resident_bytes(res, n_levels, precision="plain"):
    """Device bytes one MGPCGPoissonSolver holds: the library's pool plus the
    canonical fields of this class (b, x, x64 when not plain, and per level
    grid_type, Adiag and Ax)."""
    dim = len(res)
    canon = 0
    for l in range(n_levels):
        n = 1
        for r in level_shape(res, l):
            n *= r
        canon += n * (1 + 4 + 4 * dim)
    n0 = 1
    for r in res:
        n0 *= r
    canon += n0 * (4 + 4 + (8 if precision != "plain" else 0))
    return pool_bytes(res, n_levels, precision) + canon


class MGPCGPoissonSolver:
    """mgpcg.py:5-44.

    ``real=float`` is the reference's FP32 (Taichi default_fp).  ``precision``
    selects the arithmetic: "plain" (FP32 everywhere, as the reference),
    "mixed" (FP32 V-cycle and Laplacian, FP64 x, dot products, CG scalars and
    a true-residual recompute every ``refine_every`` iterations) or "f64".
    """

    def __init__(self, dim, res, n_mg_levels=4, pre_and_post_smoothing=2, bottom_smoothing=50,
                 real=float, precision="plain", refine_every=10, device=None):
        self.FLUID = utils.FLUID
        self.SOLID = utils.SOLID
        self.AIR = utils.AIR

        self.dim = dim
        self.res = list(res)
        self.n_mg_levels = n_mg_levels
        self.pre_and_post_smoothing = pre_and_post_smoothing
        self.bottom_smoothing = bottom_smoothing
        self.real = real
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.refine_every = refine_every
        if len(self.res) != dim:
            raise ValueError("len(res) must equal dim")
        for n in self.res:
            if n % (2 ** (n_mg_levels - 1)):
                raise ValueError(f"res {self.res} not divisible by 2^(n_mg_levels-1)")
        self.device = torch.device(device) if device is not None else torch.device("cuda")

        f32 = dict(dtype=torch.float32, device=self.device)
        self.b = torch.zeros(self.res, **f32)
        self.x = torch.zeros(self.res, **f32)
        self.grid_type = [torch.zeros(level_shape(self.res, l), dtype=torch.int8,
                                      device=self.device) for l in range(n_mg_levels)]
        self.Adiag = [torch.zeros(level_shape(self.res, l), **f32) for l in range(n_mg_levels)]
        self.Ax = [torch.zeros(level_shape(self.res, l) + [dim], **f32)
                   for l in range(n_mg_levels)]
        self.x64 = None
        self.last = None  # result dict of the last solve

        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.mgpcg_create(dim, _lib.int3(self.res), n_mg_levels,
                                          pre_and_post_smoothing, bottom_smoothing,
                                          PRECISIONS[precision], ctypes.byref(h)))
        self._h = h
        self._packed = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.mgpcg_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- assembly --------------------------------------------------------------
    def initialize(self):
        """mgpcg.py:63-77: zero b, x and every level's grid_type/Adiag/Ax."""
        self.b.zero_()
        self.x.zero_()
        for l in range(self.n_mg_levels):
            self.grid_type[l].zero_()
            self.Adiag[l].zero_()
            self.Ax[l].zero_()

    def _coarsen(self, l):
        """init_gridtype (mgpcg.py:46-60) from level l-1 into level l."""
        _lib.check(self._lib.mgpcg_coarsen(_lib.stream_handle(self.device), self.dim,
                                           _lib.int3(level_shape(self.res, l - 1)),
                                           _lib.ptr(self.grid_type[l - 1]),
                                           _lib.ptr(self.grid_type[l])))

    def reinitialize(self, cell_type, strategy):
        """mgpcg.py:79-87.  ``cell_type`` is an int tensor of shape res."""
        self._packed = False
        if not getattr(strategy, "writes_all_cells", False):
            self.initialize()
        self.grid_type[0].copy_(cell_type.to(self.grid_type[0].device))  # i32 -> i8
        strategy.build_b(self)
        strategy.build_A(self, 0)
        for l in range(1, self.n_mg_levels):
            self._coarsen(l)
            strategy.build_A(self, l)

    def reinitialize_packed(self, cell_type, strategy):
        """reinitialize for a strategy with ``assemble_packed`` (the pressure
        systems): b and every level's operator go straight into the device
        hierarchy, the same bits as reinitialize followed by load.  The
        canonical ``b``, ``Adiag`` and ``Ax`` are left as they were; the
        coarse levels' ``grid_type`` is written.  A later reinitialize or
        full_reinitialize returns to the canonical route."""
        self.grid_type[0].copy_(cell_type.to(self.grid_type[0].device))  # i32 -> i8
        strategy.assemble_packed(self)
        self._packed = True

    def full_reinitialize(self, strategy):
        """mgpcg.py:89-94: every level all FLUID."""
        self._packed = False
        if not getattr(strategy, "writes_all_cells", False):
            self.initialize()
        strategy.build_b(self)
        for l in range(self.n_mg_levels):
            self.grid_type[l].fill_(utils.FLUID)
            strategy.build_A(self, l)

    def load(self):
        """Hand the canonical fields to the device hierarchy (nothing to do
        after reinitialize_packed, which wrote the hierarchy itself)."""
        if self._packed:
            return
        h = self._h
        _lib.check(self._lib.mgpcg_set_stream(h, _lib.stream_handle(self.device)))
        for l in range(self.n_mg_levels):
            _lib.check(self._lib.mgpcg_load_level(h, l, _lib.ptr(self.grid_type[l]),
                                                  _lib.ptr(self.Adiag[l]), _lib.ptr(self.Ax[l])))
        _lib.check(self._lib.mgpcg_load_rhs(h, _lib.ptr(self.b)))

    # -- solve -------------------------------------------------------------------
    def solve(self, max_iters=-1, verbose=False, rel_tol=1e-12, abs_tol=1e-14, eps=1e-12):
        """mgpcg.py:150-209; the result is left in ``self.x`` (and ``self.x64``).

        The loop runs on the device, so ``verbose`` prints a summary after the
        solve instead of one line per iteration (mgpcg.py:161-162, 191-192).
        """
        t0 = time.perf_counter()
        self.load()
        out = _lib.MgpcgResult()
        _lib.check(self._lib.mgpcg_solve(self._h, int(max_iters), float(rel_tol), float(abs_tol),
                                         float(eps), int(self.refine_every), ctypes.byref(out)))
        self._store_x()
        self.last = out.as_dict()
        self.last["seconds"] = time.perf_counter() - t0
        if verbose:
            print(f"init rtr = {self.last['r0r0']}")
            print(f"iters {self.last['iters']}, |residual|_2={self.last['rr'] ** 0.5}, "
                  f"status {self.last['status']}")

    def solve_hostsync(self, max_iters=-1, rel_tol=1e-12, abs_tol=1e-14, eps=1e-12):
        """The same solve driven like the reference's host loop (3 syncs per iteration)."""
        self.load()
        out = _lib.MgpcgResult()
        _lib.check(self._lib.mgpcg_solve_hostsync(self._h, int(max_iters), float(rel_tol),
                                                  float(abs_tol), float(eps),
                                                  int(self.refine_every), ctypes.byref(out)))
        self._store_x()
        self.last = out.as_dict()
        return self.last

    def solve_reference_structured(self, max_iters=-1, rel_tol=1e-12, abs_tol=1e-14, eps=1e-12):
        """The plain FP32 GPU port structured like mgpcg.py (one launch per
        reference kernel, scalars on the host): the measured 1x baseline."""
        self.load()
        out = _lib.MgpcgResult()
        _lib.check(self._lib.mgpcg_solve_reference_structured(
            self._h, int(max_iters), float(rel_tol), float(abs_tol), float(eps),
            ctypes.byref(out)))
        self._store_x()
        self.last = out.as_dict()
        return self.last

    def _store_x(self):
        _lib.check(self._lib.mgpcg_set_stream(self._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_store_x(self._h, _lib.ptr(self.x), None))
        if self.precision != "plain":
            if self.x64 is None:
                self.x64 = torch.zeros(self.res, dtype=torch.float64, device=self.device)
            _lib.check(self._lib.mgpcg_store_x(self._h, None, _lib.ptr(self.x64)))

    def set_option(self, option, value):
        _lib.check(self._lib.mgpcg_set_option(self._h, option, value))

    def graph_counts(self):
        """(iteration graphs captured, iterations launched from a graph) by this
        solver's own engine (MGPCG_OPT_GRAPH)."""
        return _lib.graph_counts(solver=self._h)

    def info(self, level=0):
        out = (ctypes.c_int32 * 6)()
        nbytes = ctypes.c_int64()
        _lib.check(self._lib.mgpcg_info(self._h, level, out, ctypes.byref(nbytes)))
        return {"P0": out[0], "P1": out[1], "W": out[2], "slots": out[3], "cells": out[4],
                "coarse_from": out[5], "device_bytes": nbytes.value}

    def level_coupling(self, level):
        """(uniform, coef) of a loaded level: whether every nonzero off-diagonal
        coefficient has one value, which the kernels then use with per-cell flags."""
        u, c = ctypes.c_int(), ctypes.c_float()
        _lib.check(self._lib.mgpcg_level_coupling(self._h, level, ctypes.byref(u), ctypes.byref(c)))
        return bool(u.value), c.value

    def level_flags(self, level):
        """(flags, wet) of a loaded level as numpy arrays: the per-slot flag words
        [colour][slot] (uint16) and the per-block FLUID flags (uint8) of the
        padded colour-split layout."""
        n = self.info(level)["slots"]
        flags = torch.empty(2 * n, dtype=torch.int16, device=self.device)
        wet = torch.empty((n + 255) // 256, dtype=torch.uint8, device=self.device)
        _lib.check(self._lib.mgpcg_set_stream(self._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_level_flags(self._h, level, _lib.ptr(flags), _lib.ptr(wet)))
        return (flags.cpu().numpy().view("uint16").reshape(2, n), wet.cpu().numpy())

    def level_order(self, level):
        """(ord1, ord2) of a loaded level as int64 numpy arrays (the int32 entries:
        block index in the low 31 bits, bit 31 set for a block with a FLUID slot)
        per block of 256 slots and per half-sweep virtual block of 512 slots;
        ord2 is None except on level 0 when the batched kernels solve it."""
        inf = self.info(level)
        n = inf["slots"]
        n1 = (n + 255) // 256
        n2 = (n + 511) // 512
        o1 = torch.empty(n1, dtype=torch.int32, device=self.device)
        o2 = torch.empty(n2, dtype=torch.int32, device=self.device)
        _lib.check(self._lib.mgpcg_set_stream(self._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_level_order(self._h, level, _lib.ptr(o1), _lib.ptr(o2)))
        o2 = o2.cpu().numpy().view("uint32").astype("int64")
        return (o1.cpu().numpy().view("uint32").astype("int64"),
                o2 if level == 0 and inf["coarse_from"] > 0 else None)

    # -- kernel-level access (parity tests) --------------------------------------
    def _vec_dtype(self, field):
        if self.precision == "f64" or (field == _lib.FIELD_X and self.precision != "plain"):
            return torch.float64
        return torch.float32

    def debug_set(self, field, level, t):
        t = t.to(device=self.device, dtype=self._vec_dtype(field)).contiguous()
        _lib.check(self._lib.mgpcg_set_stream(self._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_debug_set(self._h, field, level, _lib.ptr(t)))
        _lib.synchronize(self.device)

    def debug_get(self, field, level):
        out = torch.zeros(level_shape(self.res, level), dtype=self._vec_dtype(field),
                          device=self.device)
        _lib.check(self._lib.mgpcg_set_stream(self._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_debug_get(self._h, field, level, _lib.ptr(out)))
        _lib.synchronize(self.device)
        return out

    def debug_op(self, op, level=0, arg=0):
        _lib.check(self._lib.mgpcg_set_stream(self._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_debug_op(self._h, op, level, arg))

    def time_op(self, op, level=0, arg=0, reps=50):
        """Average milliseconds of one launch of a stage (HIP events on the handle's stream)."""
        ms = ctypes.c_double()
        _lib.check(self._lib.mgpcg_set_stream(self._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_time_op(self._h, op, level, arg, reps, ctypes.byref(ms)))
        return ms.value
