"""Batched solves: many MGPCG systems in one set of persistent launches.

Every system is an assembled ``MGPCGPoissonSolver`` (same dim, precision,
n_mg_levels and smoothing counts; resolutions may differ).  ``solve_batch``
runs all of them together through mgpcg_batch_solve: each stage of an
iteration is one kernel launch over every unfinished system, systems retire
as they converge, and each result is bit-identical to solving that system
alone (the per-system arithmetic does not depend on the batch).
"""
import ctypes

from . import _lib


class BatchSolver:
    def __init__(self, device=None):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.mgpcg_batch_create(ctypes.byref(h)))
        self._h = h
        self.device = device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.mgpcg_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def graph_counts(self):
        """(iteration graphs captured, iterations launched from a graph)."""
        return _lib.graph_counts(batch=self._h)

    def solve(self, solvers, max_iters=500, rel_tol=1e-12, abs_tol=1e-14, eps=1e-12,
              refine_every=None):
        """Solve every system in ``solvers``; returns the list of result dicts."""
        if not solvers:
            return []
        refine = solvers[0].refine_every if refine_every is None else refine_every
        for s in solvers:
            s.load()
        _lib.check(self._lib.mgpcg_batch_set_stream(self._h,
                                                     _lib.stream_handle(solvers[0].device)))
        n = len(solvers)
        handles = (ctypes.c_void_p * n)(*[s._h.value for s in solvers])
        results = (_lib.MgpcgResult * n)()
        _lib.check(self._lib.mgpcg_batch_solve(self._h, handles, n, int(max_iters),
                                               float(rel_tol), float(abs_tol), float(eps),
                                               int(refine), results))
        out = []
        for s, r in zip(solvers, results):
            s._store_x()
            s.last = r.as_dict()
            out.append(s.last)
        return out


def solve_batch(solvers, **kw):
    return BatchSolver().solve(solvers, **kw)
