"""Iterations replayed from a captured hipGraph (MGPCG_OPT_GRAPH) compute the
bits the kernel-by-kernel launches compute.

The engine captures one iteration of an unchanged member set and replays it;
every replay reads its per-system scalars (iteration parity, alpha, beta,
done) from the device.  Each check stops solves after a few iterations as well
as at convergence, so the state after every replay count below is compared
with the eagerly launched solve, with the program's own eager work and
synchronizations between the replays.

`dev` = "cuda:0" runs on the MI355X (marked gpu); `dev` = "cpu" runs the CPU
emulation build, whose stream capture records the launches and replays them.
"""
import numpy as np
import pytest
import torch

from mgpcg_amd import scenes
from test_solve import SYSTEMS, _emu, _x, set_knobs


def _graph(on):
    from mgpcg_amd import _lib

    _lib.check(_lib.load().mgpcg_set_option(None, _lib.OPT_GRAPH, int(on)))


def _solver(sysd, dev, precision):
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], sysd["nu"], sysd["bottom"],
                                 precision=precision, device=dev)
    systems.assemble(s, systems.upload(sysd, dev))
    return s


def _eager_work(dev):
    """Unrelated eager device work and a synchronization between replays."""
    a = torch.arange(1 << 16, dtype=torch.float32, device=dev)
    assert float((a * 2.0).sum()) == float(a.sum()) * 2.0
    if dev != "cpu":
        torch.cuda.synchronize(dev)


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_single_solve_replays_equal_eager(dev, precision):
    from mgpcg_amd import _lib

    if _emu(dev) and precision == "mixed":
        pytest.skip("the emulator covers PLAIN; the server and batch cases cover MIXED")
    sysd = SYSTEMS["c2_3d_small"](_emu(dev))
    print(f"case {sysd['name']} {precision} graph on {dev}", flush=True)
    stops = [3, 500] if _emu(dev) else [2, 3, 5, 9, 500]
    eager = {}
    try:
        _graph(False)
        for k in stops:
            s = _solver(sysd, dev, precision)
            s.solve(k, rel_tol=1e-12)
            _lib.synchronize(dev)
            eager[k] = (_x(s), dict(s.last))
        _graph(True)
        s = _solver(sysd, dev, precision)
        for k in stops:
            _eager_work(dev)
            s.solve(k, rel_tol=1e-12)  # reuses the solver's engine
            _lib.synchronize(dev)
            x, out = _x(s), dict(s.last)
            xe, oute = eager[k]
            assert out["iters"] == oute["iters"] and out["status"] == oute["status"], (k, out)
            assert np.array_equal(x, xe), k
        captures, launches = s.graph_counts()
    finally:
        _graph(False)
    # one capture per solve (its layout after the pre-loop iteration); every
    # loop iteration after the first of a solve comes from the graph, and so
    # does the one the host queues before it reads that the solve converged
    assert captures == len(stops)
    assert launches >= sum(eager[k][1]["iters"] - 1 for k in stops)


def test_server_replays_equal_alone(dev):
    """Continuous batching with graphs: the member set changes at every
    admission and retirement (a new capture), and stays unchanged in between
    (replays)."""
    from mgpcg_amd import serve
    from mgpcg_amd import systems

    emu = _emu(dev)
    nreq = 3 if emu else 6
    lo, hi = (16, 24) if emu else (32, 64)
    reqs = [scenes.c5_request(2000 + i, lo=lo, hi=hi, L=3) for i in range(nreq)]
    try:
        _graph(True)
        set_knobs(server_depth=2, server_streams=2)
        srv = serve.Server(hbm_budget_bytes=4 << 30, precision="mixed", n_levels=3,
                           max_active=3, device=dev)
        devs = [systems.upload(r, dev) for r in reqs]
        done, i, step = [], 0, 0
        while len(done) < len(reqs):
            if i < len(reqs) and step % 3 == 0:  # staggered admission
                srv.submit(devs[i], meta=i)
                i += 1
            step += 1
            for rq in srv.step():
                rq["x"] = _x(rq["solver"])
                srv.retire(rq)
                done.append(rq)
            _eager_work(dev)
        captures, launches = srv.graph_counts()
    finally:
        _graph(False)
    assert captures > 0 and launches > captures
    for rq in done:
        sysd = reqs[rq["meta"]]
        print(f"case {sysd['name']} served with graphs", flush=True)
        s = _solver(sysd, dev, "mixed")
        s.solve(500, rel_tol=1e-12)
        xa = _x(s)
        assert s.last["iters"] == rq["result"]["iters"]
        assert np.array_equal(xa, rq["x"])


def test_batch_replays_equal_alone(dev):
    from mgpcg_amd import batch

    emu = _emu(dev)
    count, div, L = (2, 4, 3) if emu else (6, 2, 4)
    mix = scenes.c4_batch(first_seed=300, count=count, L=L)
    for sysd in mix:
        sysd.update(scenes.pool_sphere_system([n // div for n in sysd["res"]], sysd["seed"], L=L,
                                              kind=sysd["kind"]))
    solvers = [_solver(sysd, dev, "mixed") for sysd in mix]
    b = batch.BatchSolver()
    try:
        _graph(True)
        res = b.solve(solvers, max_iters=500, rel_tol=1e-12)
        xs = [_x(s) for s in solvers]
        captures, launches = b.graph_counts()
    finally:
        _graph(False)
    assert captures > 0 and launches > 0
    for sysd, xb, rb in zip(mix, xs, res):
        s = _solver(sysd, dev, "mixed")
        s.solve(500, rel_tol=1e-12)
        assert s.last["iters"] == rb["iters"] and s.last["status"] == rb["status"]
        assert np.array_equal(_x(s), xb)
