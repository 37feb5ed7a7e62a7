"""Benchmark: MGPCG solves/s with p50/p99 latency at 1e-6 relative residual.

Workload (default): the 3D 128^3 ghost-fluid pressure system -- pool y < 0.3
plus a sphere c=(0.5,0.65,0.5) r=0.15, theta >= 0.01 (c_max = 100), RHS the
divergence of a seeded U[-1,1] MAC velocity, 5 levels, 2+2 red-black sweeps,
10 bottom sweeps, rel_tol 1e-12 on squared norms (1e-6 relative residual).
Scenes (4 seeds) are uploaded to HBM before timing.

One step = one solve request: assemble b and the multigrid hierarchy on the
device from the scene and solve it to tolerance.  Requests are served with
continuous batching (serve.Server) in a closed loop that keeps `--concurrency`
requests in flight; `value` is requests completed per second over the timed
steps, p50/p99 the request latencies (submit to result).

Also reported, on the same requests:
  plain_fp32_gpu_solves_per_s   the plain FP32 GPU port structured like
                                mgpcg.py (one launch per reference kernel,
                                scalars on the host) -- the 1x baseline;
  single_request                concurrency 1 (latency-bound) throughput;
  error_vs_fp64                 relative L2 error of the mixed result against
                                an FP64 MGPCG solve to 1e-12 relative residual;
  roofline                      the level-0 red-black half-sweep, its
                                algorithmic bytes over its launch durations
                                measured on the device (first workgroup start
                                to last workgroup end) during the timed steps;
  roofline_isolated             the same with one engine stream, where no
                                other lane's kernels share the device (when
                                the headline runs more than one);
  two_lanes                     the headline workload over two engine streams
                                (when the headline runs one);
  c3_256                        the 256^3 workload, 2 in flight, one lane,
                                with `lanes`: the same over --c3-c5-streams;
  c4_batch64, c5_poisson        64 mixed systems in one batch; Poisson
                                arrivals at 4 rates on --c3-c5-streams lanes;
  roofline_dense                the same kernel on a dense 256^3 system (every
                                block has FLUID cells): its bandwidth when it
                                cannot skip work;
  cpu_baseline                  the oracle's C restatement of mgpcg.py on the
                                host cores (OpenMP), a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "taichi-ferrofluid_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell of one red-black half-sweep in colour-split FP32
# storage: own r, Adiag, type (9 B / 2), other-colour z read (4 / 2), own z
# written (4 / 2), every Ax (4 * dim)  ->  20.5 B in 3D
SMOOTH_BYTES_PER_CELL_3D = 20.5
# The reference's half-sweep (mgpcg.py:104-109) reads the type of every cell
# of its colour and does the rest only on FLUID cells: per cell of the level,
# 0.5 B (the type byte of the half of the cells in the colour) for every
# cell, and the other 20 B of the model for FLUID cells only.  The roofline
# counts FLUID cells on the device, per launch.
SMOOTH_TYPE_BYTES_PER_CELL = 0.5
SMOOTH_FLUID_BYTES_PER_CELL_3D = SMOOTH_BYTES_PER_CELL_3D - SMOOTH_TYPE_BYTES_PER_CELL
# update_xr + r.r (mgpcg.py:226-232, 211-216): gt, p, Ap, x, r read, x, r
# written (25 B) and the dot's two vectors and type (9 B), on FLUID cells; the
# type byte of every other cell
XR_FLUID_BYTES_PER_CELL = 34.0
XR_TYPE_BYTES_PER_CELL = 1.0
# The kernel moves fewer bytes than that model: on a level with uniform
# couplings it reads a 16-bit flag word (couplings, diagonal and the FLUID
# bit) instead of Ax, Adiag and the cell type (7 B per cell: own r and flag
# word 6 / 2, other-colour z 4 / 2, own z written 4 / 2), and it skips blocks
# of 256 slots without a FLUID cell.  Its physical HBM bytes per cell are
# measured with rocprofv3 counters (the committed traffic summary) and reported
# beside the model's.  On a dense system (every block FLUID) it moves 7 B per
# cell.
SMOOTH_BYTES_PER_CELL_3D_DENSE = 7.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64, help="timed solve requests")
    ap.add_argument("--warmup", type=int, default=8, help="untimed solve requests")
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS),
                    help="c2: 128^3 pool + sphere (the headline); c3: 256^3 pool + 8 spheres")
    ap.add_argument("--n", type=int, default=None, help="grid side (default: the workload's)")
    ap.add_argument("--levels", type=int, default=None)
    ap.add_argument("--seeds", type=int, default=None, help="distinct scenes cycled through")
    ap.add_argument("--concurrency", type=int, default=None, help="requests in flight")
    ap.add_argument("--precision", default="mixed", choices=["plain", "mixed", "f64"])
    ap.add_argument("--admit-per-step", type=int, default=0,
                    help="admissions per server step (0: no limit)")
    ap.add_argument("--server-depth", type=int, default=1,
                    help="iterations the server keeps queued on the device ahead of the host")
    ap.add_argument("--dense-only", action="store_true",
                    help="time only the level-0 half-sweep on the dense 256^3 system "
                         "(after the calibration copies) and print its line")
    ap.add_argument("--c3-c5-streams", dest="c5_streams", type=int, default=2,
                    help="engine lanes of the C5 server and of C3's second measurement")
    ap.add_argument("--server-streams", type=int, default=1,
                    help="engine streams the server spreads the requests in flight over")
    ap.add_argument("--canonical-assembly", action="store_true",
                    help="assemble pressure systems through the canonical fields and load "
                         "them (the reference's route) instead of straight into the packed "
                         "hierarchy")
    ap.add_argument("--baseline-steps", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mix", action="store_true",
                    help="skip the 64-system mixed batch and the Poisson-arrival sweep")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only (used for profiling runs)")
    ap.add_argument("--probe-every", type=int, default=4,
                    help="time one level-0 sweep launch in this many (1 for counter runs)")
    ap.add_argument("--probe-warmup", action="store_true",
                    help="count the warm-up requests' level-0 sweeps too (counter runs, which "
                         "attribute every dispatch of the process)")
    ap.add_argument("--autotune", action="store_true",
                    help="first tune the schedule knobs on this workload (autotune.py) and "
                         "measure with the tuned settings")
    ap.add_argument("--copy-calibration-mib", type=int, default=256,
                    help="first run a plain copy kernel over this many MiB (3 launches of each "
                         "access width), the known-size reference for HBM byte counters; 0: "
                         "none")
    a = ap.parse_args()
    w = WORKLOADS[a.workload]
    for k in ("n", "levels", "seeds", "concurrency"):
        if getattr(a, k) is None:
            setattr(a, k, w[k])
    return a


# Synthetic ghost-fluid pressure scenes (seeded, scenes.py): the 128^3 pool +
# one sphere and the 256^3 pool + 8 spheres, theta >= 0.01, U[-1, 1] MAC
# velocities.
WORKLOADS = {
    "c2": {"n": 128, "levels": 5, "seeds": 4, "concurrency": 8, "first_seed": 1,
           "scene": "c2_3d", "what": "pool+sphere"},
    "c3": {"n": 256, "levels": 6, "seeds": 2, "concurrency": 2, "first_seed": 2,
           "scene": "c3_256", "what": "pool+8 spheres"},
}


def make_scene(workload, i, n, levels):
    from mgpcg_amd import scenes

    w = WORKLOADS[workload]
    return getattr(scenes, w["scene"])(seed=w["first_seed"] + i, n=n, L=levels)


def traffic_file(workload, n):
    return os.path.join(ROOT, "profiles", f"level0_smooth_{workload}_{n}_traffic.txt")


def measured_traffic_per_cell(workload, n):
    """HBM bytes per swept cell of the level-0 half-sweep, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when it is absent)."""
    try:
        with open(traffic_file(workload, n)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_cell:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    k = (len(v) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def cpu_baseline(workload, n, levels, budget_s):
    """The oracle's C restatement of mgpcg.py (PLAIN) on every host core."""
    from oracle import core
    from oracle import systems as osys

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    sysd = make_scene(workload, 0, n, levels)
    h, b = osys.assemble(sysd)
    core.solve(h, b, max_iters=500, rel_tol=1e-12)  # warm-up
    t0 = time.perf_counter()
    count = 0
    while True:
        core.solve(h, b, max_iters=500, rel_tol=1e-12)
        count += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(count / dt, 3), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{count} solves of the {sysd['name']} system (seed {sysd['seed']}; solve "
                      f"only, assembly excluded), {dt:.1f} s, PLAIN FP32, OpenMP "
                      f"threads={threads}"}


def measure_served(args, workload, n, L, seeds, concurrency, steps, warmup, dev, streams=None,
                   rel_tol=1e-12):
    """Closed-loop continuous batching of `steps` requests with `concurrency`
    in flight; returns the throughput, latencies and the level-0 half-sweep
    roofline measured over the timed requests.  `streams` overrides the
    server's engine streams; `rel_tol` is on squared norms (1e-12: 1e-6)."""
    from mgpcg_amd import _lib
    from mgpcg_amd import serve
    from mgpcg_amd import systems
    import torch

    scenes_dev = [systems.upload(make_scene(workload, i, n, L), dev) for i in range(seeds)]
    torch.cuda.synchronize()
    lib = _lib.load()
    if streams is not None:
        _lib.check(lib.mgpcg_set_option(None, _lib.OPT_SERVER_STREAMS, streams))
    srv = serve.Server(64 << 30, precision=args.precision, n_levels=L, rel_tol=rel_tol,
                       admit_per_step=args.admit_per_step or None,
                       packed=not args.canonical_assembly, device=dev)
    if streams is not None:
        _lib.check(lib.mgpcg_set_option(None, _lib.OPT_SERVER_STREAMS, args.server_streams))
    mk = (lambda i: scenes_dev[i % len(scenes_dev)])
    if args.probe_warmup:
        srv.probe(True)
    serve.closed_loop(srv, mk, warmup, concurrency)  # allocate hierarchies, warm caches
    if not args.probe_warmup:
        srv.probe(True)
    recs, secs = serve.closed_loop(srv, mk, steps, concurrency)
    nl, ms_tot, cells_tot, fluid_tot = srv.probe(False, fluid=True)
    for r in recs:
        assert r["result"]["status"] == _lib.CONVERGED, r["result"]
    del srv
    out = {"value": steps / secs, "secs": secs, "lat": serve.latency_stats(recs),
           "iters": sorted(set(r["result"]["iters"] for r in recs)), "roofline": None,
           "scenes": scenes_dev}
    if nl > 0 and cells_tot > 0:
        out["roofline"] = smooth_roofline(nl, ms_tot, cells_tot, fluid_tot,
                                          measured_traffic_per_cell(workload, n))
    return out


def smooth_roofline(nl, ms_tot, cells_tot, fluid_tot, tpc):
    """Roofline of the probed level-0 half-sweep launches: algorithmic bytes
    of the reference's half-sweep over the cells the launches swept (20 B per
    FLUID cell + the type byte of the colour's cells, counted on the device)
    over the launch's span on the device clock (first workgroup start to last
    workgroup end); traffic from the committed counters."""
    cells = cells_tot / nl
    fluid = fluid_tot / nl
    alg_bytes = SMOOTH_FLUID_BYTES_PER_CELL_3D * fluid + SMOOTH_TYPE_BYTES_PER_CELL * cells
    launch_s = ms_tot * 1e-3 / nl
    achieved = alg_bytes / launch_s / 1e9
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": round(tpc * cells) if tpc is not None else None,
        "kernel": "mg::k_stage<float, 3, 0, 2> (smooth(0, 0), the level-0 red-black "
                  "half-sweep over every member system, probed launches)",
        "alg_model": "20 B per level-0 FLUID cell + 0.5 B per level-0 cell",
        "alg_bytes_per_launch": round(alg_bytes), "launch_us": round(launch_s * 1e6, 3),
        "launches": nl, "cells_per_launch": round(cells), "fluid_per_launch": round(fluid),
        "cells_total": cells_tot, "fluid_total": fluid_tot,
        "traffic_bytes_per_cell": tpc,
        # measured HBM bytes of a launch over its duration: the bandwidth the
        # kernel actually draws
        "achieved_traffic": (round(tpc * cells / launch_s / 1e9, 1) if tpc is not None
                             else None),
        "frac_traffic": (round(tpc * cells / launch_s / 1e9 / HBM_PEAK_GBS, 4)
                         if tpc is not None else None)}


def fp64_error(args, scene_dev, n, L, dev, rel_tols=(1e-12, 1e-14)):
    """Relative L2 error of a result of the precision under test against FP64
    MGPCG converged to 1e-12 relative residual (rel_tol 1e-24 on squared
    norms), for each stopping tolerance in `rel_tols` (squared norms: 1e-12
    is the reference's 1e-6 relative residual)."""
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems
    import torch

    s64 = mgpcg.MGPCGPoissonSolver(3, [n] * 3, L, 2, 10, precision="f64", device=dev)
    systems.assemble(s64, scene_dev)
    s64.solve(500, rel_tol=1e-12)  # FP64 MGPCG stopped at the 1e-6 residual
    x64s = s64.x64.clone()
    it64s = s64.last["iters"]
    s64.solve(4000, rel_tol=1e-24)
    x64 = s64.x64.clone()
    out = {"fp64_iters": s64.last["iters"],
           "fp64_rel_residual": (s64.last["rr"] / s64.last["r0r0"]) ** 0.5,
           "fp64_same_tol_iters": it64s}
    del s64
    sm = mgpcg.MGPCGPoissonSolver(3, [n] * 3, L, 2, 10, precision=args.precision, device=dev)
    for tol in rel_tols:
        systems.assemble(sm, scene_dev)
        sm.solve(500, rel_tol=tol)
        xm = (sm.x64 if sm.x64 is not None else sm.x).double()
        rec = {"rel_residual_target": tol ** 0.5,
               "rel_l2": float(torch.linalg.norm(xm - x64) / torch.linalg.norm(x64)),
               "iters": sm.last["iters"],
               "true_rel_residual": ((sm.last["true_rr"] / sm.last["r0r0"]) ** 0.5
                                     if sm.last["true_rr"] >= 0 else None)}
        if tol == 1e-12:
            rec["rel_l2_vs_fp64_mgpcg_same_tol"] = float(torch.linalg.norm(xm - x64s) /
                                                         torch.linalg.norm(x64s))
            out.update(rec)
        else:
            out[f"at_rel_residual_{tol ** 0.5:.0e}"] = rec
    del sm
    return out


def measure_dense(args, dev, n=256, levels=6, reps=100):
    """The level-0 half-sweep on one dense system: 256^3, every interior cell
    FLUID (a 1-cell SOLID border), no ghost-fluid terms, so every block has work
    and every diagonal is a plain sum.  Launch time from HIP events around
    `reps` back-to-back launches (the gaps between them included)."""
    from mgpcg_amd import _lib
    from mgpcg_amd import mgpcg
    import numpy as np
    from mgpcg_amd import scenes
    from mgpcg_amd import systems
    import torch

    res = [n] * 3
    dx = 1.0 / n
    phi = np.full(res, -1.0, np.float32)
    sysd = scenes._pressure(res, dx, phi, scenes.cell_types(phi),
                            scenes.mac_velocity(res, np.random.default_rng(5)), levels, False,
                            100.0, 1e-3, 5, f"dense_{n}")
    s = mgpcg.MGPCGPoissonSolver(3, res, levels, 2, 10, precision=args.precision, device=dev)
    systems.assemble(s, systems.upload(sysd, dev))
    s.load()
    torch.cuda.synchronize()
    ms = s.time_op(_lib.OP_SMOOTH, 0, 0, reps=reps)
    cells = n ** 3
    launch_s = ms * 1e-3
    model = SMOOTH_BYTES_PER_CELL_3D * cells / launch_s / 1e9
    own = SMOOTH_BYTES_PER_CELL_3D_DENSE * cells / launch_s / 1e9
    del s
    # On this system the kernel reads 16-bit flag words where the reference
    # model reads FP32 couplings and diagonals: its own 7 B per cell are the
    # bytes it moves, and the frac; the 20.5 B model's figure is beside it.
    return {"system": f"{n}^3, every interior cell FLUID, uniform couplings, plain-sum diagonals",
            "kernel": "mg::k_stage<float, 3, 0, 1>", "launch_us": round(ms * 1e3, 3),
            "cells_per_launch": cells, "bytes_per_cell": SMOOTH_BYTES_PER_CELL_3D_DENSE,
            "achieved": round(own, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(own / HBM_PEAK_GBS, 4),
            "model_bytes_per_cell": SMOOTH_BYTES_PER_CELL_3D,
            "achieved_model": round(model, 1)}


def measure_batch64(args, dev):
    """64 uneven systems (sides in {32, 48, 64, 96, 128} per axis) mixing the
    four per-substep kinds (psi, pressure, surface tension, volume control),
    assembled on the device and solved together in one batch."""
    from mgpcg_amd import batch
    from mgpcg_amd import mgpcg
    from mgpcg_amd import scenes
    from mgpcg_amd import systems
    import torch

    L = 5
    mix = scenes.c4_batch(first_seed=100, count=64, L=L)
    devs = [systems.upload(sd, dev) for sd in mix]
    solvers = [mgpcg.MGPCGPoissonSolver(3, sd["res"], L, 2, 10, precision=args.precision,
                                        device=dev) for sd in mix]
    bs = batch.BatchSolver(dev)

    def run():
        for s, d in zip(solvers, devs):
            systems.assemble(s, d)
        return bs.solve(solvers, max_iters=500, rel_tol=1e-12)

    run()  # warm-up
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        res = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    its = [r["iters"] for r in res]
    conv = sum(r["status"] == 1 for r in res)
    capped = [sd["name"] for sd, r in zip(mix, res) if r["status"] == 2]
    assert conv + len(capped) == len(mix), [r["status"] for r in res]
    return {"solves_per_s": round(len(mix) / dt, 3), "ms_per_batch": round(1e3 * dt, 3),
            "cells": int(sum(int(np.prod(sd["res"])) for sd in mix)),
            "iters_min_max": [min(its), max(its)],
            "converged": conv, "converged_solves_per_s": round(conv / dt, 3),
            "hit_max_iters": capped,
            "what": "assembly + solve of 64 systems per batch, rel residual 1e-6, 500 "
                    "iterations at most (a system at the cap counts in solves_per_s, not in "
                    "converged_solves_per_s)"}


def measure_poisson(args, dev, rates=(50, 200, 1000, 2000), per_rate=48, pool=16):
    """Open-loop serving: requests arrive at Poisson times with rate lambda,
    sides log-uniform in [64, 256] per axis (pool + sphere scenes)."""
    from mgpcg_amd import _lib
    from mgpcg_amd import scenes
    from mgpcg_amd import serve
    from mgpcg_amd import systems
    import torch

    L = 5
    scs = [systems.upload(scenes.c5_request(1000 + i, L=L), dev) for i in range(pool)]
    torch.cuda.synchronize()
    # the lane count is read when the server is made
    _lib.check(_lib.load().mgpcg_set_option(None, _lib.OPT_SERVER_STREAMS, args.c5_streams))
    srv = serve.Server(96 << 30, precision=args.precision, n_levels=L, device=dev)
    _lib.check(_lib.load().mgpcg_set_option(None, _lib.OPT_SERVER_STREAMS, args.server_streams))
    serve.closed_loop(srv, lambda i: scs[i % pool], pool, 8)  # resident hierarchies, warm
    out = []
    for lam in rates:
        recs, secs = serve.poisson_arrivals(srv, scs, lam, per_rate, seed=lam)
        for r in recs:
            assert r["result"]["status"] == 1, r["result"]
        st = serve.latency_stats(recs)
        out.append({"lambda": lam, "solves_per_s": round(len(recs) / secs, 3),
                    "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                    "converged": len(recs), "max_members": srv.max_members,
                    "server_streams": args.c5_streams})
    del srv
    return out


def profile_summary(db_path, bench_path, command):
    """Text summary of a rocprofv3 kernel trace (its rocpd SQLite output) of a
    bench run, beside that run's JSON line: the probed level-0 half-sweep's
    rocprofv3 durations for exactly the launches the bench timed (the first
    `launches` dispatches of its probed instantiation: the headline runs
    first), the kernel-time breakdown inside the headline's timed window, and
    roofline / latency rows for update_xr and the coarse-level kernel."""
    import sqlite3

    line = None
    with open(bench_path) as f:
        for ln in f:
            ln = ln.strip()
            if ln.startswith("{") and '"metric"' in ln:
                line = json.loads(ln)
    assert line is not None, "no bench JSON line in " + bench_path
    rf = line["roofline"]
    con = sqlite3.connect(db_path)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    probed = [r for r in rows if "k_stage<float, 3, 0, 2>" in r[0]]
    head = probed[:rf["launches"]]
    assert len(head) == rf["launches"], (len(head), rf["launches"])
    t0, t1 = head[0][1], head[-1][2]
    avg_us = sum(r[2] - r[1] for r in head) / len(head) / 1e3
    win = [r for r in rows if r[1] >= t0 and r[2] <= t1]
    agg = {}
    def short_name(name):
        # drop the return type and the parameter list (the final balanced "(...)")
        name = name[5:] if name.startswith("void ") else name
        if name.endswith(")"):
            depth = 0
            for i in range(len(name) - 1, -1, -1):
                depth += {")": 1, "(": -1}.get(name[i], 0)
                if depth == 0:
                    return name[:i]
        return name

    for name, a, b in win:
        short = short_name(name)
        c = agg.setdefault(short, [0, 0.0])
        c[0] += 1
        c[1] += (b - a) / 1e3
    total = sum(v[1] for v in agg.values())
    out = []
    w = out.append
    w(f"# rocprofv3 kernel trace of the served workload: {line['config']['workload']}")
    w("")
    w(f"command (from the repository root, on the GPU box, after "
      f"`cd /tmp && export TMPDIR=/tmp` in the shell):")
    w("")
    w(f"    {command}")
    w("")
    w(f"bench line of that run: value {line['value']} {line['unit']}, p50 {line['p50_ms']} ms, "
      f"p99 {line['p99_ms']} ms, iters {line['iters']}, steps {line['steps']}, "
      f"warmup {line['warmup']}")
    w("")
    frac_prof = rf["alg_bytes_per_launch"] / (avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS
    w("dominant bandwidth kernel: the level-0 red-black half-sweep smooth(0, 0), probed launches")
    w(f"  instantiation                    mg::k_stage<float, 3, 0, 2>")
    w(f"  launches the bench timed         {rf['launches']} (first {len(head)} dispatches of "
      f"that name in the trace)")
    w(f"  rocprofv3 average duration       {avg_us:.3f} us")
    w(f"  bench device-clock span average  {rf['launch_us']} us (first workgroup start to "
      f"last workgroup end)")
    w(f"  difference                       {avg_us - rf['launch_us']:.3f} us per launch: "
      f"the profiler's interval adds the dispatch before the first workgroup and the "
      f"completion signal after the last")
    w(f"  level-0 cells / FLUID per launch {rf['cells_per_launch']} / {rf['fluid_per_launch']}")
    w(f"  algorithmic bytes per launch     {rf['alg_bytes_per_launch']} ({rf['alg_model']})")
    w(f"  frac of 8 TB/s, rocprofv3 time   {frac_prof:.4f}")
    w(f"  frac of 8 TB/s, bench (roofline) {rf['frac']}")
    cells, fluid = rf["cells_per_launch"], rf["fluid_per_launch"]
    xr = next((k for k in agg if k.startswith("mg::k_cg_xr<")), None)
    if xr:
        n_xr, t_xr = agg[xr]
        xr_bytes = XR_FLUID_BYTES_PER_CELL * fluid + XR_TYPE_BYTES_PER_CELL * (cells - fluid)
        xr_avg = t_xr / n_xr
        w("")
        w(f"update_xr + r.r ({xr}):")
        w(f"  launches in the window           {n_xr}, average {xr_avg:.3f} us, "
          f"{100 * t_xr / total:.1f}% of kernel time")
        w(f"  algorithmic bytes per launch     {round(xr_bytes)} (34 B per level-0 FLUID cell "
          f"+ 1 B per other cell; cells as the half-sweep's member set)")
        w(f"  frac of 8 TB/s                   "
          f"{xr_bytes / (xr_avg * 1e-6) / 1e9 / HBM_PEAK_GBS:.4f}")
    co = next((k for k in agg if k.startswith("mg::k_coarse<")), None)
    if co:
        n_co, t_co = agg[co]
        w("")
        w(f"coarse levels ({co}): latency bound, one 1024-lane workgroup per member system")
        w(f"  launches in the window           {n_co}, average {t_co / n_co:.3f} us, "
          f"{100 * t_co / total:.1f}% of kernel time")
    w("")
    w(f"kernel time inside the headline's timed window ({(t1 - t0) / 1e6:.3f} ms of wall "
      f"time, {total / 1e3:.3f} ms of kernel time on all streams):")
    w("")
    w(f"{'kernel':<96} {'calls':>6} {'total_us':>10} {'avg_us':>9} {'pct':>6}")
    for k, (n_, t_) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:24]:
        w(f"{k[:96]:<96} {n_:>6} {t_:>10.1f} {t_ / n_:>9.3f} {100 * t_ / total:>6.2f}")
    return "\n".join(out)


def traffic_summary(fetch_db, write_db, fetch_log, write_log, command):
    """Text summary of the two counter runs (FETCH_SIZE, then WRITE_SIZE, in
    KiB per dispatch) of a bench run made with --probe-warmup --probe-every 1
    --copy-calibration-mib M: the probed level-0 half-sweep's HBM bytes per
    swept level-0 cell, each counter scaled by its reading for the known-size
    calibration copies (4-byte accesses, the solver kernels' width)."""
    import sqlite3

    def bench_line(path):
        with open(path) as f:
            for ln in f:
                ln = ln.strip()
                if ln.startswith("{") and '"metric"' in ln:
                    return json.loads(ln)
        raise AssertionError("no bench JSON line in " + path)

    def counter(db, name):
        con = sqlite3.connect(db)
        return con.execute("select name, counter_value from pmc_events where counter_name = ? "
                           "order by dispatch_id", (name,)).fetchall()

    out = []
    w = out.append
    w("two counter runs of the same bench command (one counter each), from the repository")
    w("root on the GPU box after `cd /tmp && export TMPDIR=/tmp` in the shell:")
    w("")
    wl = command.split("--workload", 1)[1].split()[0] if "--workload" in command else "c2"
    tag = wl if "--workload" in command else "head"  # the driver's headline command
    where = (f"in the headline command ({wl} workload, " if tag == "head"
             else f", {wl} workload (")
    out.insert(0, f"# HBM traffic of the level-0 half-sweep {where}"
                  f"{WORKLOADS[wl]['n']}^3 {WORKLOADS[wl]['what']})".replace(" , ", ", "))
    out.insert(1, "")
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        w(f"    timeout -s KILL 900 rocprofv3 --pmc {ctr} -d gpurun_out/pmc_{tag}_{ctr} -o {wl} -- "
          f"python3 bench.py {command}")
    w("")
    w("The bench probes one dispatch in --probe-every of the level-0 half-sweep smooth(0, 0)")
    w("(instantiation mg::k_stage<float, 3, 0, 2>) inside the headline's timed region (with")
    w("--probe-warmup from the warm-up on) and counts, on the device, the level-0 cells and")
    w("FLUID cells those dispatches swept. They are the first `launches` dispatches of that")
    w("instantiation in the process, so the counters are summed over exactly those.")
    w("")
    res = {}
    for ctr, db, log in (("FETCH_SIZE", fetch_db, fetch_log), ("WRITE_SIZE", write_db, write_log)):
        rows = counter(db, ctr)
        line = bench_line(log)
        rf = line["roofline"]
        cal_v = [v for n_, v in rows if "MulFunctor" in n_][:3]
        cal_s = [v for n_, v in rows if "CUDAFunctorOnSelf_add" in n_][:3]
        mib = line["config"]["copy_calibration_mib"]
        kib = mib * 1024.0
        sv = kib / (sum(cal_v) / len(cal_v))
        ss = kib / (sum(cal_s) / len(cal_s))
        # the bench's cell counts cover its first `launches` probed dispatches
        # (the headline's timed region runs first; with --probe-warmup its
        # warm-up too)
        k = [v for n_, v in rows if "k_stage<float, 3, 0, 2>" in n_][:rf["launches"]]
        assert len(k) == rf["launches"], (len(k), rf["launches"])
        res[ctr] = {"scale_vec": sv, "scale_scalar": ss, "dispatches": len(k),
                    "launches": rf["launches"], "cells": rf["cells_total"],
                    "fluid": rf["fluid_total"], "kib": sum(k)}
    w(f"calibration copies, {mib} MiB read + {mib} MiB written, 3 launches each:")
    for lab, key in (("vector (16-byte accesses)", "scale_vec"),
                     ("scalar (4-byte accesses)", "scale_scalar")):
        w(f"  {lab:<27} FETCH_SIZE scale {res['FETCH_SIZE'][key]:.4f};  "
          f"WRITE_SIZE scale {res['WRITE_SIZE'][key]:.4f}")
    w("")
    f_, wr = res["FETCH_SIZE"], res["WRITE_SIZE"]
    fetched = f_["kib"] * 1024 * f_["scale_scalar"] / f_["cells"]
    written = wr["kib"] * 1024 * wr["scale_scalar"] / wr["cells"]
    w("kernel mg::k_stage<float, 3, 0, 2> (level-0 half-sweep, probed, 4-byte accesses):")
    w(f"  dispatches (fetch run / write run)   {f_['dispatches']} / {wr['dispatches']}")
    w(f"  bench-probed launches                {f_['launches']} / {wr['launches']}")
    w(f"  level-0 cells swept                  {f_['cells']} / {wr['cells']}")
    w(f"  FLUID cells swept                    {f_['fluid']} / {wr['fluid']}")
    w(f"  raw FETCH_SIZE bytes per cell        {f_['kib'] * 1024 / f_['cells']:.3f}")
    w(f"  raw WRITE_SIZE bytes per cell        {wr['kib'] * 1024 / wr['cells']:.3f}")
    w(f"  fetched per cell, scalar scale       {fetched:.3f}")
    w(f"  written per cell, scalar scale       {written:.3f}")
    w("  model bytes per cell                 20 per FLUID cell + 0.5 per cell (the")
    w("                                       reference's half-sweep); the kernel reads")
    w("                                       16-bit flag words on uniform-coupling levels")
    w("                                       and skips blocks of 256 slots without FLUID")
    w("")
    w(f"traffic_bytes_per_cell: {fetched + written:.3f}")
    return "\n".join(out)


def dense_summary(fetch_db, write_db, trace_db, log):
    """Text summary of three rocprofv3 runs of `bench.py --dense-only` (a kernel
    trace, then FETCH_SIZE, then WRITE_SIZE in KiB per dispatch): the dense
    256^3 half-sweep's HBM bytes per cell (each counter scaled by its reading
    for the known-size calibration copies, 4-byte accesses) and its rocprofv3
    durations, beside the line's HIP-event launch time."""
    import sqlite3

    line = None
    with open(log) as f:
        for ln in f:
            ln = ln.strip()
            if ln.startswith("{") and '"roofline_dense"' in ln:
                line = json.loads(ln)
    assert line is not None, "no --dense-only line in " + log
    rd = line["roofline_dense"]
    kname = "k_stage<float, 3, 0, 1>"

    def counter(db, name):
        con = sqlite3.connect(db)
        return con.execute("select name, counter_value from pmc_events where counter_name = ? "
                           "order by dispatch_id", (name,)).fetchall()

    res = {}
    kib_cal = line["copy_calibration_mib"] * 1024.0
    for ctr, db in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        rows = counter(db, ctr)
        cal_s = [v for n_, v in rows if "CUDAFunctorOnSelf_add" in n_][:3]
        k = [v for n_, v in rows if kname in n_]
        assert cal_s and k, (ctr, len(cal_s), len(k))
        res[ctr] = {"scale": kib_cal / (sum(cal_s) / len(cal_s)), "n": len(k),
                    "kib": sum(k) / len(k)}
    con = sqlite3.connect(trace_db)
    durs = [(b - a) / 1e3 for name, a, b in
            con.execute("select name, start, end from kernels order by start").fetchall()
            if kname in name]
    assert durs
    avg_us = sum(durs) / len(durs)
    cells = rd["cells_per_launch"]
    fetched = res["FETCH_SIZE"]["kib"] * 1024 * res["FETCH_SIZE"]["scale"]
    written = res["WRITE_SIZE"]["kib"] * 1024 * res["WRITE_SIZE"]["scale"]
    own = rd["bytes_per_cell"] * cells
    out = []
    w = out.append
    w(f"# HBM traffic and duration of the level-0 half-sweep on the dense system ({rd['system']})")
    w("")
    w("three runs of the same command, from the repository root on the GPU box after")
    w("`cd /tmp && export TMPDIR=/tmp` in the shell:")
    w("")
    w("    rocprofv3 --kernel-trace --stats -d gpurun_out/dense_trace -o dense -- "
      "python3 bench.py --dense-only")
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        w(f"    timeout -s KILL 300 rocprofv3 --pmc {ctr} -d gpurun_out/dense_{ctr} -o dense -- "
          f"python3 bench.py --dense-only")
    w("")
    w(f"Every dispatch of mg::{kname} in that process is the dense system's half-sweep")
    w("smooth(0, 0) (the warm-up and the timed launches of its HIP-event timing).")
    w("")
    w(f"calibration copies: {line['copy_calibration_mib']} MiB, 4-byte accesses: FETCH_SIZE scale "
      f"{res['FETCH_SIZE']['scale']:.4f}, WRITE_SIZE scale {res['WRITE_SIZE']['scale']:.4f}")
    w("")
    w(f"kernel mg::{kname}, {cells} level-0 cells (one colour swept per launch):")
    w(f"  dispatches (trace / fetch run / write run)  {len(durs)} / {res['FETCH_SIZE']['n']} / "
      f"{res['WRITE_SIZE']['n']}")
    w(f"  rocprofv3 average duration                  {avg_us:.3f} us")
    w(f"  HIP-event launch time of the bench line     {rd['launch_us']} us (back-to-back launches)")
    w(f"  fetched per cell, scaled                    {fetched / cells:.3f} B")
    w(f"  written per cell, scaled                    {written / cells:.3f} B")
    w(f"  traffic per cell                            {(fetched + written) / cells:.3f} B "
      f"(the kernel's own model: {rd['bytes_per_cell']} B)")
    tb = (fetched + written) / (avg_us * 1e-6) / 1e12
    w(f"  traffic / rocprofv3 duration                {tb:.3f} TB/s, frac {tb * 1e3 / HBM_PEAK_GBS:.4f}")
    ob = own / (avg_us * 1e-6) / 1e12
    w(f"  own bytes / rocprofv3 duration              {ob:.3f} TB/s, frac {ob * 1e3 / HBM_PEAK_GBS:.4f}")
    w(f"  own bytes / HIP-event launch time           {rd['achieved'] / 1e3:.3f} TB/s, "
      f"frac {rd['frac']}")
    return "\n".join(out)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--traffic-summary":
        # bench.py --traffic-summary <fetch db> <write db> <fetch log> <write log> <bench args>
        print(traffic_summary(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5],
                              " ".join(sys.argv[6:])))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--dense-summary":
        # bench.py --dense-summary <fetch db> <write db> <trace db> <dense-only log>
        print(dense_summary(*sys.argv[2:6]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--profile-summary":
        # bench.py --profile-summary <rocpd db> <file with the bench JSON line> <command>
        print(profile_summary(sys.argv[2], sys.argv[3], " ".join(sys.argv[4:])))
        return
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1"}))
        print("bench.py: only --gpus 1 is supported: the solver runs on one GPU (how it would "
              "shard is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    from mgpcg_amd import _lib
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.copy_calibration_mib:
        # two known-size copies: 16-byte vector accesses (aligned views) and
        # 4-byte scalar accesses (views offset by one element, which the
        # elementwise kernel cannot vectorize) -- the access width of the
        # solver's kernels
        nel = args.copy_calibration_mib << 18
        src = torch.ones(nel + 1, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src[:nel], 1.0, out=dst[:nel])
        for _ in range(3):
            torch.add(src[1:], 0.0, out=dst[1:])
        torch.cuda.synchronize()
        del src, dst
    _lib.check(_lib.load().mgpcg_set_option(None, _lib.OPT_PROBE_EVERY, args.probe_every))
    _lib.check(_lib.load().mgpcg_set_option(None, _lib.OPT_SERVER_DEPTH, args.server_depth))
    _lib.check(_lib.load().mgpcg_set_option(None, _lib.OPT_SERVER_STREAMS, args.server_streams))
    if args.dense_only:
        print(json.dumps({"roofline_dense": measure_dense(args, dev),
                          "copy_calibration_mib": args.copy_calibration_mib}), flush=True)
        return
    wl, n, L = args.workload, args.n, args.levels
    what = WORKLOADS[wl]["what"]

    tuned = None
    if args.autotune:
        from mgpcg_amd import autotune

        scs = [systems.upload(make_scene(wl, i, n, L), dev) for i in range(args.seeds)]
        tuned = autotune.tune(lambda i: scs[i % len(scs)], L, concurrency=args.concurrency,
                              precision=args.precision, device=dev,
                              log=lambda t: print("autotune", json.dumps(t), file=sys.stderr,
                                                  flush=True))
        args.server_depth = tuned["knobs"]["server_depth"]
        args.server_streams = tuned["knobs"]["server_streams"]
        del scs

    # headline: continuous batching, closed loop
    head = measure_served(args, wl, n, L, args.seeds, args.concurrency, args.steps, args.warmup,
                          dev)
    scenes_dev = head.pop("scenes")
    line = {
        "metric": "MGPCG solves/s + p50/p99 latency at 1e-6 rel. residual, 128^3-256^3, 1 GPU",
        "value": round(head["value"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * head["secs"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"plain": "f32", "mixed": "f32+f64", "f64": "f64"}[args.precision],
        "data": "synthetic (seeded scenes, resident in HBM)",
        "config": {"workload": f"3D {n}^3 ghost-fluid {what} pressure solve requests, "
                               f"L={L}, 2+2 sweeps, bottom 10, rel residual 1e-6, "
                               f"{args.precision} precision, continuous batching with "
                               f"{args.concurrency} requests in flight",
                   "n": n, "levels": L, "seeds": args.seeds, "concurrency": args.concurrency,
                   "admit_per_step": args.admit_per_step, "server_depth": args.server_depth,
                   "server_streams": args.server_streams,
                   "assembly": "canonical" if args.canonical_assembly else "packed",
                   "copy_calibration_mib": args.copy_calibration_mib},
        **({"autotune": {"knobs": tuned["knobs"], "solves_per_s": tuned["solves_per_s"],
                         "start": tuned["start"], "trials": len(tuned["trials"])}}
           if tuned else {}),
        "p50_ms": round(head["lat"]["p50_ms"], 3),
        "p99_ms": round(head["lat"]["p99_ms"], 3),
        "iters": head["iters"],
        "roofline": head["roofline"],
        "cpu_baseline": None,
    }

    if not args.no_extras:
        if args.server_streams > 1:
            # the same kernel with no other lane's kernels beside it on the
            # device: the headline's launches share it with the second lane
            # warm-up of 2 x concurrency requests: the new server's first
            # requests create their hierarchies (mgpcg_create synchronizes)
            iso = measure_served(args, wl, n, L, args.seeds, args.concurrency,
                                 max(32, args.steps // 2), 2 * args.concurrency, dev,
                                 streams=1)
            iso.pop("scenes")
            line["roofline_isolated"] = dict(iso["roofline"] or {}, server_streams=1,
                                             solves_per_s=round(iso["value"], 3))
        else:
            # the same requests spread over two engine streams: one lane's
            # latency-bound kernels beside the other's level-0 kernels
            # (warm-up of 2 x concurrency requests, as above)
            two = measure_served(args, wl, n, L, args.seeds, args.concurrency,
                                 max(32, args.steps // 2), 2 * args.concurrency, dev,
                                 streams=2)
            two.pop("scenes")
            line["two_lanes"] = {"solves_per_s": round(two["value"], 3),
                                 "p50_ms": round(two["lat"]["p50_ms"], 3),
                                 "p99_ms": round(two["lat"]["p99_ms"], 3),
                                 "roofline": two["roofline"]}
        # single request in flight: the latency-bound rate
        one = measure_served(args, wl, n, L, args.seeds, 1, max(4, args.steps // 8), 2, dev)
        one.pop("scenes")
        line["single_request"] = {"solves_per_s": round(one["value"], 3),
                                  **{k: round(v, 3) for k, v in one["lat"].items()}}
        # the plain FP32 GPU port structured like mgpcg.py (the 1x baseline)
        s = mgpcg.MGPCGPoissonSolver(3, [n] * 3, L, 2, 10, precision="plain", device=dev)
        for i in range(2):  # warm-up
            systems.assemble(s, scenes_dev[i % len(scenes_dev)])
            s.solve_reference_structured(500, rel_tol=1e-12)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it_p = []
        for i in range(args.baseline_steps):
            systems.assemble(s, scenes_dev[i % len(scenes_dev)])
            out = s.solve_reference_structured(500, rel_tol=1e-12)
            assert out["status"] == _lib.CONVERGED, out
            it_p.append(out["iters"])
        torch.cuda.synchronize()
        base = args.baseline_steps / (time.perf_counter() - t0)
        line["plain_fp32_gpu_solves_per_s"] = round(base, 3)
        line["plain_fp32_gpu_iters"] = sorted(set(it_p))
        line["speedup_vs_plain_fp32_gpu"] = round(head["value"] / base, 3)
        del s
        # error of a result against FP64 MGPCG converged to 1e-12 relative
        # residual, at the reference's 1e-6 residual and at 1e-7 (where the
        # error falls under 1e-5), with the served rate at 1e-7
        line["error_vs_fp64"] = fp64_error(args, scenes_dev[0], n, L, dev)
        tight = measure_served(args, wl, n, L, args.seeds, args.concurrency,
                               max(16, args.steps), 4, dev, rel_tol=1e-14)
        tight.pop("scenes")
        line["error_vs_fp64"]["at_rel_residual_1e-07"]["served_solves_per_s"] = \
            round(tight["value"], 3)
        line["error_vs_fp64"]["at_rel_residual_1e-07"]["served_iters"] = tight["iters"]
        del scenes_dev
        if wl == "c2":
            # the 256^3 workload, where a system's level 0 outgrows the Infinity Cache
            c3 = WORKLOADS["c3"]
            big = measure_served(args, "c3", c3["n"], c3["levels"], c3["seeds"],
                                 c3["concurrency"], 8, 2, dev)
            big_scenes = big.pop("scenes")
            line["c3_256"] = {"solves_per_s": round(big["value"], 3),
                              "p50_ms": round(big["lat"]["p50_ms"], 3),
                              "p99_ms": round(big["lat"]["p99_ms"], 3), "iters": big["iters"],
                              "concurrency": c3["concurrency"], "roofline": big["roofline"],
                              "error_vs_fp64": fp64_error(args, big_scenes[0], c3["n"],
                                                          c3["levels"], dev)}
            del big_scenes
            # one system per lane: one lane's coarse levels beside the other's level 0
            # (its roofline launch shares the device, so the line above keeps its own)
            if args.c5_streams != args.server_streams:
                big2 = measure_served(args, "c3", c3["n"], c3["levels"], c3["seeds"],
                                      c3["concurrency"], 8, 2, dev, streams=args.c5_streams)
                big2.pop("scenes")
                line["c3_256"]["lanes"] = {"server_streams": args.c5_streams,
                                           "solves_per_s": round(big2["value"], 3),
                                           "p50_ms": round(big2["lat"]["p50_ms"], 3),
                                           "p99_ms": round(big2["lat"]["p99_ms"], 3)}
            line["roofline_dense"] = measure_dense(args, dev)
            if not args.no_mix:
                line["c4_batch64"] = measure_batch64(args, dev)
                line["c5_poisson"] = measure_poisson(args, dev)
        if not args.no_cpu:
            line["cpu_baseline"] = cpu_baseline(wl, n, L, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
