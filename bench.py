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
                                other lane's kernels share the device;
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
# The kernel moves fewer bytes than that model: on a level with uniform
# couplings it reads a 16-bit flag word (couplings and diagonal) instead of Ax
# and Adiag (7.5 B per cell: own r, type and flag word 7 / 2, other-colour z
# 4 / 2, own z written 4 / 2), and it skips blocks of 256 slots without a
# FLUID cell.  Its physical HBM bytes per cell are measured with rocprofv3
# counters (the committed traffic summary) and reported beside the model's.
# On a dense system (every block FLUID) it moves 7.5 B per cell.
SMOOTH_BYTES_PER_CELL_3D_DENSE = 7.5


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
    ap.add_argument("--server-depth", type=int, default=2,
                    help="iterations the server keeps queued on the device ahead of the host")
    ap.add_argument("--server-streams", type=int, default=2,
                    help="engine streams the server spreads the requests in flight over")
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
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a plain copy kernel over this many MiB (3 launches), the "
                         "known-size reference for HBM byte counters")
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
    import scenes

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


def measure_served(args, workload, n, L, seeds, concurrency, steps, warmup, dev, streams=None):
    """Closed-loop continuous batching of `steps` requests with `concurrency`
    in flight; returns the throughput, latencies and the level-0 half-sweep
    roofline measured over the timed requests.  `streams` overrides the
    server's engine streams."""
    import _lib
    import serve
    import systems
    import torch

    scenes_dev = [systems.upload(make_scene(workload, i, n, L), dev) for i in range(seeds)]
    torch.cuda.synchronize()
    lib = _lib.load()
    if streams is not None:
        _lib.check(lib.mgpcg_set_option(None, _lib.OPT_SERVER_STREAMS, streams))
    srv = serve.Server(64 << 30, precision=args.precision, n_levels=L,
                       admit_per_step=args.admit_per_step or None, device=dev)
    if streams is not None:
        _lib.check(lib.mgpcg_set_option(None, _lib.OPT_SERVER_STREAMS, args.server_streams))
    mk = (lambda i: scenes_dev[i % len(scenes_dev)])
    if args.probe_warmup:
        srv.probe(True)
    serve.closed_loop(srv, mk, warmup, concurrency)  # allocate hierarchies, warm caches
    if not args.probe_warmup:
        srv.probe(True)
    recs, secs = serve.closed_loop(srv, mk, steps, concurrency)
    nl, ms_tot, cells_tot = srv.probe(False)
    for r in recs:
        assert r["result"]["status"] == _lib.CONVERGED, r["result"]
    del srv
    out = {"value": steps / secs, "secs": secs, "lat": serve.latency_stats(recs),
           "iters": sorted(set(r["result"]["iters"] for r in recs)), "roofline": None,
           "scenes": scenes_dev}
    if nl > 0 and cells_tot > 0:
        alg_bytes = SMOOTH_BYTES_PER_CELL_3D * cells_tot / nl
        launch_s = ms_tot * 1e-3 / nl
        achieved = alg_bytes / launch_s / 1e9
        tpc = measured_traffic_per_cell(workload, n)
        out["roofline"] = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(tpc * cells_tot / nl) if tpc is not None else None,
            "kernel": "mg::k_stage<float, 3, 0, 1> (smooth(0, 0), level-0 red-black half-sweep "
                      "over every member system)",
            "alg_bytes_per_launch": round(alg_bytes), "launch_us": round(launch_s * 1e6, 3),
            "launches": nl, "cells_per_launch": round(cells_tot / nl),
            "traffic_bytes_per_cell": tpc,
            # measured HBM bytes of a launch over its duration: the bandwidth the
            # kernel actually draws (the model's bytes include cells it skips)
            "achieved_traffic": (round(tpc * cells_tot / nl / launch_s / 1e9, 1)
                                 if tpc is not None else None),
            "frac_traffic": (round(tpc * cells_tot / nl / launch_s / 1e9 / HBM_PEAK_GBS, 4)
                             if tpc is not None else None)}
    return out


def measure_dense(args, dev, n=256, levels=6, reps=100):
    """The level-0 half-sweep on one dense system: 256^3, every interior cell
    FLUID (a 1-cell SOLID border), no ghost-fluid terms, so every block has work
    and every diagonal is a plain sum.  Launch time from HIP events around
    `reps` back-to-back launches (the gaps between them included)."""
    import _lib
    import mgpcg
    import numpy as np
    import scenes
    import systems
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
    achieved = SMOOTH_BYTES_PER_CELL_3D * cells / launch_s / 1e9
    own = SMOOTH_BYTES_PER_CELL_3D_DENSE * cells / launch_s / 1e9
    del s
    return {"system": f"{n}^3, every interior cell FLUID, uniform couplings, plain-sum diagonals",
            "kernel": "mg::k_stage<float, 3, 0, 1>", "launch_us": round(ms * 1e3, 3),
            "cells_per_launch": cells, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "kernel_bytes_per_cell": SMOOTH_BYTES_PER_CELL_3D_DENSE,
            "achieved_kernel_bytes": round(own, 1), "frac_kernel_bytes": round(own / HBM_PEAK_GBS, 4)}


def measure_batch64(args, dev):
    """64 uneven systems (sides in {32, 48, 64, 96, 128} per axis) mixing the
    four per-substep kinds (psi, pressure, surface tension, volume control),
    assembled on the device and solved together in one batch."""
    import batch
    import mgpcg
    import scenes
    import systems
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
    return {"solves_per_s": round(len(mix) / dt, 3), "ms_per_batch": round(1e3 * dt, 3),
            "cells": int(sum(int(np.prod(sd["res"])) for sd in mix)),
            "iters_min_max": [min(its), max(its)],
            "converged": sum(r["status"] == 1 for r in res),
            "what": "assembly + solve of 64 systems per batch, rel residual 1e-6"}


def measure_poisson(args, dev, rates=(50, 200, 1000, 2000), per_rate=48, pool=16):
    """Open-loop serving: requests arrive at Poisson times with rate lambda,
    sides log-uniform in [64, 256] per axis (pool + sphere scenes)."""
    import scenes
    import serve
    import systems
    import torch

    L = 5
    scs = [systems.upload(scenes.c5_request(1000 + i, L=L), dev) for i in range(pool)]
    torch.cuda.synchronize()
    srv = serve.Server(96 << 30, precision=args.precision, n_levels=L, device=dev)
    serve.closed_loop(srv, lambda i: scs[i % pool], pool, 8)  # resident hierarchies, warm
    out = []
    for lam in rates:
        recs, secs = serve.poisson_arrivals(srv, scs, lam, per_rate, seed=lam)
        st = serve.latency_stats(recs)
        out.append({"lambda": lam, "solves_per_s": round(len(recs) / secs, 3),
                    "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                    "max_members": srv.max_members})
    del srv
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1"}))
        print("bench.py: only --gpus 1 is supported: the solver runs on one GPU (how it would "
              "shard is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    import _lib
    import mgpcg
    import systems

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
    wl, n, L = args.workload, args.n, args.levels
    what = WORKLOADS[wl]["what"]

    tuned = None
    if args.autotune:
        import autotune

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
                   "server_streams": args.server_streams},
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
            iso = measure_served(args, wl, n, L, args.seeds, args.concurrency,
                                 max(16, args.steps // 4), 4, dev, streams=1)
            iso.pop("scenes")
            line["roofline_isolated"] = dict(iso["roofline"] or {}, server_streams=1,
                                             solves_per_s=round(iso["value"], 3))
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
        # error of a mixed result against FP64 MGPCG converged to 1e-12 relative residual
        sm = mgpcg.MGPCGPoissonSolver(3, [n] * 3, L, 2, 10, precision=args.precision,
                                      device=dev)
        systems.assemble(sm, scenes_dev[0])
        sm.solve(500, rel_tol=1e-12)
        xm = (sm.x64 if sm.x64 is not None else sm.x).double()
        s64 = mgpcg.MGPCGPoissonSolver(3, [n] * 3, L, 2, 10, precision="f64", device=dev)
        systems.assemble(s64, scenes_dev[0])
        s64.solve(500, rel_tol=1e-12)  # FP64 MGPCG stopped at the same 1e-6 residual
        x64s = s64.x64.clone()
        it64s = s64.last["iters"]
        err_same = float(torch.linalg.norm(xm - x64s) / torch.linalg.norm(x64s))
        s64.solve(2000, rel_tol=1e-24)
        x64 = s64.x64
        err = float(torch.linalg.norm(xm - x64) / torch.linalg.norm(x64))
        true_rel = (sm.last["true_rr"] / sm.last["r0r0"]) ** 0.5 if sm.last["true_rr"] >= 0 \
            else None
        line["error_vs_fp64"] = {"rel_l2": err,
                                 "rel_l2_vs_fp64_mgpcg_same_tol": err_same,
                                 "fp64_same_tol_iters": it64s,
                                 "fp64_iters": s64.last["iters"],
                                 "fp64_rel_residual": (s64.last["rr"] / s64.last["r0r0"]) ** 0.5,
                                 "true_rel_residual": true_rel}
        del sm, s64, scenes_dev
        if wl == "c2":
            # the 256^3 workload, where a system's level 0 outgrows the Infinity Cache
            c3 = WORKLOADS["c3"]
            big = measure_served(args, "c3", c3["n"], c3["levels"], c3["seeds"],
                                 c3["concurrency"], 8, 2, dev)
            big.pop("scenes")
            line["c3_256"] = {"solves_per_s": round(big["value"], 3),
                              "p50_ms": round(big["lat"]["p50_ms"], 3),
                              "p99_ms": round(big["lat"]["p99_ms"], 3), "iters": big["iters"],
                              "concurrency": c3["concurrency"], "roofline": big["roofline"]}
            line["roofline_dense"] = measure_dense(args, dev)
            if not args.no_mix:
                line["c4_batch64"] = measure_batch64(args, dev)
                line["c5_poisson"] = measure_poisson(args, dev)
        if not args.no_cpu:
            line["cpu_baseline"] = cpu_baseline(wl, n, L, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
