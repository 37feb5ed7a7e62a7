"""Autotuner of the solver's launch shape and smoothing schedule.

The library's schedule knobs (mgpcg_set_option) are process-wide, and none of
them changes a bit of any result: they choose how the same arithmetic is laid
out on the device.

  max_grid        workgroups of a persistent stage launch (the launch shape)
  coarse_max      levels with at most this many slots per colour run in the
                  one-workgroup coarse kernel (read when a handle is created)
  restrict_rows   restriction: 2 residual pass + child gather, 1 row-tiled
                  with the residuals in LDS, 0 one lane per child
  restrict_coarse the same on the levels below level 0 (-1: as restrict_rows)
  fuse_first      fill(0) + smooth(0) + smooth(1) in one launch (1 every
                  level, 2 the levels below level 0)
  fuse_prolong    prolongation inside the next smooth(1)
  fuse_p          update_p inside compute_Ap
  server_streams  engine lanes of a server (read when a server is created)
  server_depth    iterations a server keeps queued ahead of the host

``tune`` measures the served solves/s of the caller's own workload (closed
loop, ``concurrency`` requests in flight) for candidate settings, by
coordinate descent from the current settings: a pass tries every value of one
knob with the others held and keeps a value only when it beats the best so far
by ``min_gain``; passes repeat until one changes nothing.  Every candidate's
results must equal the first measurement's bit for bit, or the tuner raises.
The winner is left applied (``restore=True`` puts the starting settings back).
"""
import ctypes
import json
import os
import time

from . import _lib

OPTIONS = {
    "max_grid": _lib.OPT_MAX_GRID,
    "coarse_max": _lib.OPT_COARSE_MAX_SLOTS,
    "restrict_rows": _lib.OPT_RESTRICT_ROWS,
    "restrict_coarse": _lib.OPT_RESTRICT_COARSE,
    "fuse_first": _lib.OPT_FUSE_FIRST,
    "fuse_prolong": _lib.OPT_FUSE_PROLONG,
    "fuse_p": _lib.OPT_FUSE_P,
    "server_streams": _lib.OPT_SERVER_STREAMS,
    "server_depth": _lib.OPT_SERVER_DEPTH,
}

# candidate values per knob, in the order they are tried
SPACE = {
    "max_grid": [512, 1024, 2048, 4096],
    "coarse_max": [1024, 4096, 16384],
    "restrict_rows": [2, 1, 0],
    "restrict_coarse": [-1, 0, 1, 2],
    "fuse_first": [0, 1, 2],
    "fuse_prolong": [1, 0],
    "fuse_p": [0, 1],
    "server_streams": [1, 2, 3],
    "server_depth": [1, 2, 3],
}


def current(names=None):
    """The library's current value of each knob."""
    lib = _lib.load()
    out = {}
    for k in names or OPTIONS:
        v = ctypes.c_int()
        _lib.check(lib.mgpcg_get_option(None, OPTIONS[k], ctypes.byref(v)))
        out[k] = v.value
    return out


def apply(knobs):
    lib = _lib.load()
    for k, v in knobs.items():
        _lib.check(lib.mgpcg_set_option(None, OPTIONS[k], int(v)))


def _served(make_scene, n_levels, precision, concurrency, requests, warmup, device):
    """Served rate of a new server (so server- and handle-creation knobs take
    effect) in steady state: `concurrency` requests stay in flight throughout,
    and the clock runs from the `warmup`-th completion to the
    (`warmup` + `requests`)-th.  Returns (solves/s, {request index: x of the
    timed requests}); x stays on the device (a copy in stream order), so
    keeping it does not stall the host's submissions."""
    from . import serve

    srv = serve.Server(64 << 30, precision=precision, n_levels=n_levels, device=device)
    xs = {}
    issued, finished, t0, t1 = 0, 0, None, None
    total = warmup + requests
    while finished < total:
        while len(srv.running) + len(srv.queue) < concurrency:
            srv.submit(make_scene(issued), meta=issued)
            issued += 1
        for req in srv.step():
            if warmup <= finished < total:
                sv = req["solver"]
                xs[req["meta"]] = (sv.x64 if sv.x64 is not None else sv.x).clone()
            srv.retire(req)
            finished += 1
            if finished == warmup:
                t0 = time.perf_counter()
            if finished == total:
                t1 = time.perf_counter()
    while srv.running or srv.queue:  # drain the requests issued beyond the window
        for req in srv.step():
            srv.retire(req)
    _lib.synchronize(srv.device)
    del srv
    if t0 is None:  # warmup == 0
        raise ValueError("warmup must be at least 1")
    return requests / (t1 - t0), xs


def tune(make_scene, n_levels, concurrency=8, requests=48, warmup=8, precision="mixed",
         space=None, min_gain=0.01, max_passes=3, restore=False, device="cuda:0", log=None):
    """Coordinate-descent tuning of `space` (default: every knob of SPACE) on
    the workload make_scene(i) -> uploaded system (systems.upload).  Returns
    {"knobs": best settings, "solves_per_s": their rate, "start": starting
    settings and rate, "trials": every measurement}."""
    import torch

    space = dict(SPACE if space is None else space)
    start = current(list(space))
    best = dict(start)
    trials = []
    ref = {}

    def measure(knobs):
        apply(knobs)
        rate, xs = _served(make_scene, n_levels, precision, concurrency, requests, warmup,
                           device)
        for i, x in xs.items():  # request i is the same system in every trial
            if i not in ref:
                ref[i] = x
            elif not torch.equal(x, ref[i]):
                raise RuntimeError(f"autotune: settings {knobs} changed the result of request {i}")
        trials.append({"knobs": dict(knobs), "solves_per_s": round(rate, 3)})
        if log:
            log(trials[-1])
        return rate

    try:
        # the first servers of a process pay for code loading and first-touch
        # allocation: one untimed trial of the starting settings first
        apply(best)
        _served(make_scene, n_levels, precision, concurrency, concurrency, concurrency, device)
        best_rate = measure(best)
        start_rate = best_rate
        for _ in range(max_passes):
            changed = False
            for name, values in space.items():
                for v in values:
                    if v == best[name]:
                        continue
                    cand = dict(best, **{name: v})
                    rate = measure(cand)
                    if rate > best_rate * (1.0 + min_gain):
                        best, best_rate, changed = cand, rate, True
            if not changed:
                break
    finally:
        apply(start if restore else best)
    return {"knobs": best, "solves_per_s": round(best_rate, 3),
            "start": {"knobs": start, "solves_per_s": round(start_rate, 3)}, "trials": trials}


def cache_key(res, n_levels, precision, concurrency, device="cuda:0"):
    import torch

    name = (torch.cuda.get_device_name(device) if torch.device(device).type == "cuda"
            else "cpu")
    return f"{name}|{'x'.join(map(str, res))}|L{n_levels}|{precision}|c{concurrency}"


def save(path, key, result):
    """Store a tuning result under `key` in the JSON file at `path`."""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = {"knobs": result["knobs"], "solves_per_s": result["solves_per_s"]}
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def load(path, key):
    """Apply and return the knobs stored under `key`, or None when absent."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    if key not in data:
        return None
    apply(data[key]["knobs"])
    return data[key]["knobs"]
