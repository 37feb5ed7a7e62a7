"""Serving: continuous batching of MGPCG solve requests on one GPU.

A request is a scene (scenes.py) to assemble and solve.  The loop:
  * admits queued requests while the HBM budget allows, onto a resident
    hierarchy of the request's shape (hierarchies are kept on the device and
    reused; a shape that does not fit evicts idle hierarchies of other shapes);
  * assembles the admitted system (strategy kernels; pressure systems
    straight into the packed hierarchy unless ``packed=False``) on one of
    ``assembly_streams`` side streams in turn, while the engine's stream
    iterates the systems already running: assemblies overlap each other and
    the iterations, and a system joins the running ones at the first
    iteration boundary after its assembly finished;
  * calls mgpcg_server_step: the new systems join at the next iteration
    boundary, one iteration of every member is enqueued, and members the
    device reports done (one iteration behind) retire;
  * reports each request's latency (arrival to result) and the solves/s.

Each result is bit-identical to solving the same system alone: the per-
system arithmetic does not depend on who else is in the batch.
"""
import collections
import ctypes
import time

import torch

from . import _lib
from . import mgpcg
from . import scenes
from . import systems


class Server:
    def __init__(self, hbm_budget_bytes, precision="mixed", n_levels=5, nu=2, bottom=10,
                 max_iters=500, rel_tol=1e-12, abs_tol=1e-14, eps=1e-12, refine_every=10,
                 max_active=None, admit_per_step=None, packed=True, assembly_streams=8,
                 device="cuda:0"):
        self.device = torch.device(device)
        self.budget = int(hbm_budget_bytes)
        self.precision = precision
        self.L, self.nu, self.bottom = n_levels, nu, bottom
        self.refine_every = refine_every
        self.max_active = max_active
        # at most this many admissions per step: the host's assembly work then
        # fits in the iteration the device is running, and a cohort that
        # finished together re-enters staggered
        self.admit_per_step = admit_per_step
        self.packed = packed
        # default-priority side streams for assemblies (none for the CPU emulation)
        self._asm = ([torch.cuda.Stream(device=self.device) for _ in range(assembly_streams)]
                     if self.device.type == "cuda" and assembly_streams > 0 else [])
        self._asm_next = 0
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.mgpcg_server_create(int(max_iters), float(rel_tol), float(abs_tol),
                                                 float(eps), int(refine_every), ctypes.byref(h)))
        self._h = h
        self.free = collections.defaultdict(list)  # shape -> idle resident solvers
        self.lru = collections.OrderedDict()  # id(solver) -> solver (idle, for eviction)
        self.bytes = {}  # id(solver) -> device bytes
        self.used = 0
        self.running = {}  # tag -> (solver, request record)
        self.queue = collections.deque()
        self.done = []
        self.next_tag = 0
        self.max_members = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.mgpcg_server_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- resident hierarchies -----------------------------------------------------
    def _solver_bytes(self, res):
        return mgpcg.resident_bytes(list(res), self.L, self.precision)

    def _acquire(self, res):
        key = tuple(res)
        if self.free[key]:
            s = self.free[key].pop()
            self.lru.pop(id(s), None)
            return s
        need = self._solver_bytes(res)
        while self.used + need > self.budget and self.lru:
            # evict the least recently used idle hierarchy
            _, victim = self.lru.popitem(last=False)
            self.free[tuple(victim.res)].remove(victim)
            self.used -= self.bytes.pop(id(victim))
            del victim
        if self.used + need > self.budget:
            return None
        s = mgpcg.MGPCGPoissonSolver(len(res), list(res), self.L, self.nu, self.bottom,
                                     precision=self.precision, refine_every=self.refine_every,
                                     device=self.device)
        self.bytes[id(s)] = need
        self.used += need
        return s

    def _release(self, s):
        self.free[tuple(s.res)].append(s)
        self.lru[id(s)] = s

    # -- requests -------------------------------------------------------------------
    def submit(self, scene_dev, arrival=None, meta=None):
        """Queue an uploaded scene; returns its tag."""
        tag = self.next_tag
        self.next_tag += 1
        self.queue.append({"tag": tag, "scene": scene_dev,
                           "arrival": time.perf_counter() if arrival is None else arrival,
                           "meta": meta})
        return tag

    def _admit(self):
        admitted = 0
        ready = None  # the server's stream so far: scene uploads, the last readbacks
        while self.queue:
            if self.max_active is not None and len(self.running) >= self.max_active:
                break
            if self.admit_per_step is not None and admitted >= self.admit_per_step and \
                    self.running:
                break
            req = self.queue[0]
            s = self._acquire(req["scene"]["res"])
            if s is None:
                if not self.running:
                    raise RuntimeError("request does not fit the HBM budget even alone")
                break
            self.queue.popleft()
            if self._asm:
                # the assembly reads the scene, which was uploaded on the
                # server's stream (one event per admission round)
                if ready is None:
                    ready = torch.cuda.Event()
                    ready.record(torch.cuda.current_stream(self.device))
                st = self._asm[self._asm_next]
                self._asm_next = (self._asm_next + 1) % len(self._asm)
                st.wait_event(ready)
                with torch.cuda.stream(st):
                    self._assemble(s, req)
            else:
                self._assemble(s, req)
            req["admitted"] = time.perf_counter()
            self.running[req["tag"]] = (s, req)
            admitted += 1
        return admitted

    def _assemble(self, s, req):
        """Assemble and submit on torch's current stream."""
        systems.assemble(s, req["scene"], packed=self.packed)
        s.load()
        _lib.check(self._lib.mgpcg_set_stream(s._h, _lib.stream_handle(self.device)))
        _lib.check(self._lib.mgpcg_server_submit(self._h, s._h, req["tag"]))

    def step(self):
        """Admit, advance every member one iteration, collect finished results."""
        self._admit()
        _lib.check(self._lib.mgpcg_server_set_stream(self._h, _lib.stream_handle(self.device)))
        nact = ctypes.c_int()
        _lib.check(self._lib.mgpcg_server_step(self._h, ctypes.byref(nact)))
        self.max_members = max(self.max_members, nact.value)
        return self._poll()

    def _poll(self):
        cap = 64
        tags = (ctypes.c_int64 * cap)()
        res = (_lib.MgpcgResult * cap)()
        n = ctypes.c_int()
        finished = []
        while True:
            _lib.check(self._lib.mgpcg_server_poll(self._h, tags, res, cap, ctypes.byref(n)))
            for i in range(n.value):
                s, req = self.running.pop(tags[i])
                s._store_x()
                s.last = res[i].as_dict()
                req["result"] = dict(s.last)
                req["finished"] = time.perf_counter()
                req["solver"] = s
                finished.append(req)
            if n.value < cap:
                break
        return finished

    def retire(self, req, keep_x=False):
        """Give a finished request's hierarchy back to the idle pool."""
        s = req.pop("solver")
        if keep_x:
            req["x"] = s.x.clone()
        self._release(s)

    def probe(self, enable, fluid=False):
        """Start (True) / stop (False) timing the level-0 half-sweep launches;
        stopping returns (launches, device ms, level-0 cells swept), plus the
        FLUID cells among them when ``fluid``."""
        n, ms, cells, fl = ctypes.c_int(), ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._lib.mgpcg_server_probe(self._h, int(bool(enable)), ctypes.byref(n),
                                                ctypes.byref(ms), ctypes.byref(cells),
                                                ctypes.byref(fl)))
        if fluid:
            return n.value, ms.value, cells.value, fl.value
        return n.value, ms.value, cells.value

    def graph_counts(self):
        """(iteration graphs captured, iterations launched from a graph), summed
        over the engine streams."""
        return _lib.graph_counts(server=self._h)

    def busy(self):
        npend, nact, infl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.mgpcg_server_pending(self._h, ctypes.byref(npend),
                                                  ctypes.byref(nact), ctypes.byref(infl)))
        return bool(self.queue or self.running or npend.value or nact.value or infl.value)


def closed_loop(server, make_scene, total, concurrency, keep_x=False):
    """Keep `concurrency` requests in flight until `total` have finished.

    Returns (records, seconds).
    """
    t0 = time.perf_counter()
    issued = 0
    records = []
    while len(records) < total:
        while issued < total and len(server.running) + len(server.queue) < concurrency:
            server.submit(make_scene(issued))
            issued += 1
        for req in server.step():
            server.retire(req, keep_x=keep_x)
            records.append(req)
    _lib.synchronize(server.device)
    return records, time.perf_counter() - t0


def poisson_arrivals(server, scenes_dev, rate, total, seed=0, keep_x=False):
    """Open loop: request i arrives at a Poisson(rate) time; returns (records, seconds)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, total)
    arrive = np.cumsum(gaps)
    t0 = time.perf_counter()
    issued = 0
    records = []
    while len(records) < total:
        now = time.perf_counter() - t0
        while issued < total and arrive[issued] <= now:
            server.submit(scenes_dev[issued % len(scenes_dev)], arrival=t0 + arrive[issued])
            issued += 1
        if server.running or server.queue:
            for req in server.step():
                server.retire(req, keep_x=keep_x)
                records.append(req)
        elif issued < total:
            time.sleep(max(0.0, min(arrive[issued] - now, 0.001)))
    _lib.synchronize(server.device)
    return records, time.perf_counter() - t0


def latency_stats(records):
    lat = sorted(r["finished"] - r["arrival"] for r in records)

    def q(p):
        k = (len(lat) - 1) * p
        lo = int(k)
        hi = min(lo + 1, len(lat) - 1)
        return lat[lo] + (lat[hi] - lat[lo]) * (k - lo)

    return {"p50_ms": 1e3 * q(0.5), "p99_ms": 1e3 * q(0.99), "mean_ms": 1e3 * sum(lat) / len(lat)}
