"""Benchmark: MGPCG solves/s with p50/p99 latency at 1e-6 relative residual.

Workload (default): the 3D 128^3 ghost-fluid pressure system -- pool y < 0.3
plus a sphere c=(0.5,0.65,0.5) r=0.15, theta >= 0.01, seeded U[-1,1] MAC
velocity divergence as RHS, 5 levels, 2+2 red-black sweeps, 10 bottom
sweeps, rel_tol 1e-12 on squared norms.  One step = one whole solve request:
assemble b and the hierarchy on the device from the scene fields (already in
HBM) and solve to tolerance.  The JSON line carries, besides the headline
mixed-precision rate:
  * the plain FP32 GPU MGPCG driven like the reference (host round trip after
    every reduction) and the device-resident plain FP32 loop;
  * roofline: the level-0 red-black half-sweep (the dominant kernel), its
    algorithmic bytes per launch over its HIP-event-timed launch duration;
  * cpu_baseline: the C restatement of mgpcg.py on the host cores (OpenMP).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "taichi-ferrofluid_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell of one red-black half-sweep in colour-split FP32
# storage: own r, Adiag, type (9 B / 2), other-colour z read (4 / 2), own z
# written (4 / 2), every Ax (4 * dim)  ->  20.5 B in 3D
SMOOTH_BYTES_PER_CELL_3D = 20.5


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128, help="grid side of the workload")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--seeds", type=int, default=4, help="distinct scenes cycled through")
    ap.add_argument("--precision", default="mixed", choices=["plain", "mixed", "f64"])
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-baselines", action="store_true")
    return ap.parse_args()


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    k = (len(v) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def cpu_baseline(n, levels, budget_s):
    """The oracle's C restatement of mgpcg.py (PLAIN) on every host core."""
    import numpy as np  # noqa: F401

    import scenes
    from oracle import core
    from oracle import systems as osys

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    sysd = scenes.c2_3d(seed=1, n=n, L=levels)
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
    return {"value": count / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{count} solves of the same {n}^3 system (solve only, assembly excluded), "
                      f"{dt:.1f} s, PLAIN FP32, OMP threads={threads}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1"}))
        print("bench.py: only --gpus 1 is supported (single-GPU solver; multi-GPU sharding "
              "is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    import _lib
    import mgpcg
    import scenes
    import systems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    n, L = args.n, args.levels
    scenes_dev = [systems.upload(scenes.c2_3d(seed=1 + i, n=n, L=L), dev)
                  for i in range(args.seeds)]
    torch.cuda.synchronize()

    def run(precision, steps, warmup, hostsync=False):
        solver = mgpcg.MGPCGPoissonSolver(3, [n] * 3, L, 2, 10, precision=precision, device=dev)
        lat, iters, solve_only = [], [], []
        for i in range(warmup + steps):
            sc = scenes_dev[i % len(scenes_dev)]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            systems.assemble(solver, sc)
            t1 = time.perf_counter()
            if hostsync:
                solver.solve_hostsync(500, rel_tol=1e-12)
            else:
                solver.solve(500, rel_tol=1e-12)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if i >= warmup:
                lat.append(t2 - t0)
                solve_only.append(t2 - t1)
                iters.append(solver.last["iters"])
                assert solver.last["status"] == _lib.CONVERGED, solver.last
        return solver, lat, iters, solve_only

    # headline
    torch.cuda.synchronize()
    solver, lat, iters, solve_only = run(args.precision, args.steps, args.warmup)
    total = sum(lat)
    value = args.steps / total

    extra = {}
    if not args.no_baselines:
        _, lat_p, it_p, _ = run("plain", args.plain_steps, 2, hostsync=True)
        _, lat_d, it_d, _ = run("plain", args.plain_steps, 2)
        extra["plain_fp32_hostsync_solves_per_s"] = args.plain_steps / sum(lat_p)
        extra["plain_fp32_device_loop_solves_per_s"] = args.plain_steps / sum(lat_d)
        extra["speedup_vs_plain_hostsync"] = value / extra["plain_fp32_hostsync_solves_per_s"]
        extra["iters_plain"] = sorted(set(it_p))

    # roofline of the dominant kernel: level-0 half-sweep, timed live with HIP events
    ms = solver.time_op(_lib.OP_SMOOTH, 0, 0, reps=100)
    cells = n ** 3
    alg_bytes = SMOOTH_BYTES_PER_CELL_3D * cells
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "mg::k_stage<float, 3, 0> (smooth(0, 0): level-0 red-black half-sweep)",
                "alg_bytes_per_launch": alg_bytes, "launch_us": round(ms * 1e3, 3)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, L, args.cpu_seconds)

    line = {
        "metric": "MGPCG solves/s + p50/p99 latency at 1e-6 rel. residual, 128^3-256^3, 1 GPU",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"plain": "f32", "mixed": "f32+f64", "f64": "f64"}[args.precision],
        "data": "synthetic (seeded scenes, resident in HBM)",
        "config": {"workload": f"3D {n}^3 ghost-fluid pool+sphere pressure solve, L={L}, "
                               f"rel residual 1e-6, {args.precision} precision",
                   "n": n, "levels": L, "seeds": args.seeds},
        "p50_ms": round(1e3 * pct(lat, 0.5), 3),
        "p99_ms": round(1e3 * pct(lat, 0.99), 3),
        "solve_only_p50_ms": round(1e3 * pct(solve_only, 0.5), 3),
        "iters": sorted(set(iters)),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update({k: (round(v, 3) if isinstance(v, float) else v) for k, v in extra.items()})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
