"""Whole-solve parity of the HIP solver against the oracle.

Bar: pressures within 1e-5 relative L2 of the oracle at 1e-6 relative
residual (rel_tol 1e-12 on squared norms, mgpcg.py:164), iteration count
within +-1, same status; MIXED additionally within 1e-5 of an FP64 solve
converged to 1e-12 relative residual.  Batched and served solves must be
bit-identical to the same system solved alone.

`dev` = "cuda:0" runs on the MI355X (marked gpu) at the workload sizes;
`dev` = "cpu" runs the CPU emulation build of the same HIP sources at
reduced sizes.
"""
import numpy as np
import pytest

from mgpcg_amd import scenes
from helpers import rel_l2
from oracle import core
from oracle import systems as osys

PREC = {"plain": core.PLAIN, "mixed": core.MIXED, "f64": core.F64}


def gpu_solve(sysd, dev, precision="plain", hostsync=False, max_iters=500, rel_tol=1e-12):
    from mgpcg_amd import _lib
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    d = systems.upload(sysd, dev)
    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], sysd["nu"], sysd["bottom"],
                                 precision=precision, device=dev)
    systems.assemble(s, d)
    if hostsync:
        s.solve_hostsync(max_iters, rel_tol=rel_tol)
    else:
        s.solve(max_iters, rel_tol=rel_tol)
    _lib.synchronize(dev)
    return _x(s), dict(s.last), s


def _x(s):
    return (s.x64 if s.x64 is not None else s.x).cpu().numpy().astype(np.float64)


def _emu(dev):
    return dev == "cpu"


SYSTEMS = {
    "c1_2d": lambda emu: scenes.c1_2d(seed=0, n=64 if emu else 128),
    "c2_3d_small": lambda emu: scenes.c2_3d(seed=1, n=24 if emu else 64, L=3 if emu else 4),
    "c2_3d_128": lambda emu: None if emu else scenes.c2_3d(seed=1, n=128, L=5),
    "c3_256": lambda emu: None if emu else scenes.c3_256(seed=2, n=256, L=6),
}


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_solve_matches_oracle(dev, name, precision):
    sysd = SYSTEMS[name](_emu(dev))
    if sysd is None:
        pytest.skip("workload size runs on the GPU only")
    print(f"case {sysd['name']} {precision} on {dev}", flush=True)
    x, out, _ = gpu_solve(sysd, dev, precision)
    xr, outr, _ = osys.solve(sysd, PREC[precision], max_iters=500, rel_tol=1e-12)
    assert out["status"] == 1 == outr["status"], (out, outr)
    assert abs(out["iters"] - outr["iters"]) <= 1, (out["iters"], outr["iters"])
    assert rel_l2(x, xr) <= 1e-5
    assert out["rr"] < out["tol"]
    if precision == "plain":
        # FP32 arithmetic in the same order as the oracle, order-insensitive dots
        assert out["iters"] == outr["iters"]
        assert np.array_equal(x, xr)


def test_hostsync_loop_equals_device_loop(dev):
    sysd = SYSTEMS["c2_3d_small"](_emu(dev))
    xa, outa, _ = gpu_solve(sysd, dev, "plain")
    xb, outb, _ = gpu_solve(sysd, dev, "plain", hostsync=True)
    assert outa["iters"] == outb["iters"]
    assert np.array_equal(xa, xb)


def test_reference_structured_equals_device_loop(dev):
    """The plain baseline (one launch per reference kernel, scalars on the
    host) computes exactly what the batched engine computes in PLAIN."""
    from mgpcg_amd import _lib
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    sysd = SYSTEMS["c2_3d_small"](_emu(dev))
    xa, outa, _ = gpu_solve(sysd, dev, "plain")
    d = systems.upload(sysd, dev)
    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], sysd["nu"], sysd["bottom"],
                                 precision="plain", device=dev)
    systems.assemble(s, d)
    outb = s.solve_reference_structured(500, rel_tol=1e-12)
    _lib.synchronize(dev)
    assert outb["iters"] == outa["iters"] and outb["status"] == outa["status"] == 1
    assert np.array_equal(_x(s), xa)


KNOB_DEFAULTS = {"coarse_max": 4096, "max_grid": 2048, "fuse_first": 0, "fuse_prolong": 1,
                 "fuse_p": 0, "compact": 1, "coarse_lds": 1, "restrict_rows": 2,
                 "restrict_coarse": -1}


def set_knobs(**kw):
    from mgpcg_amd import _lib

    lib = _lib.load()
    opt = {"coarse_max": _lib.OPT_COARSE_MAX_SLOTS, "max_grid": _lib.OPT_MAX_GRID,
           "fuse_first": _lib.OPT_FUSE_FIRST, "fuse_prolong": _lib.OPT_FUSE_PROLONG,
           "fuse_p": _lib.OPT_FUSE_P, "compact": _lib.OPT_COMPACT,
           "coarse_lds": _lib.OPT_COARSE_LDS, "restrict_rows": _lib.OPT_RESTRICT_ROWS,
           "restrict_coarse": _lib.OPT_RESTRICT_COARSE,
           "probe_every": _lib.OPT_PROBE_EVERY, "server_depth": _lib.OPT_SERVER_DEPTH,
           "server_streams": _lib.OPT_SERVER_STREAMS}
    for k, v in kw.items():
        _lib.check(lib.mgpcg_set_option(None, opt[k], v))


@pytest.mark.parametrize("nu", [1, 3])
@pytest.mark.parametrize("fuse_first", [0, 1])
def test_smoothing_counts_match_oracle(dev, nu, fuse_first):
    """Smoothing counts other than 2 per side (which half-sweep forms the
    colour-1 residual, and whether level 0's first half-sweep comes from the
    kernel that wrote r, depend on them and on fusing fill(0) + smooth(0) +
    smooth(1)): PLAIN results are the oracle's bits."""
    sysd = dict(SYSTEMS["c2_3d_small"](_emu(dev)), nu=nu)
    print(f"case {sysd['name']} nu={nu} fuse_first={fuse_first} on {dev}", flush=True)
    set_knobs(fuse_first=fuse_first)
    try:
        x, out, _ = gpu_solve(sysd, dev, "plain")
    finally:
        set_knobs(**KNOB_DEFAULTS)
    xr, outr, _ = osys.solve(sysd, core.PLAIN, max_iters=500, rel_tol=1e-12)
    assert out["status"] == 1 == outr["status"], (out, outr)
    assert out["iters"] == outr["iters"]
    assert np.array_equal(x, xr)


# single-knob settings that a combined setting below also exercises (the
# emulator runs only the combined one; the GPU runs both)
EMU_COVERED = [{"max_grid": 64}, {"fuse_first": 1}, {"fuse_p": 1}, {"coarse_max": 1024},
               {"compact": 0}, {"coarse_lds": 0}, {"restrict_rows": 1}, {"restrict_rows": 0},
               {"fuse_first": 2}, {"restrict_coarse": 0}]


@pytest.mark.parametrize("knobs", [
    {"coarse_max": 0}, {"coarse_max": 1 << 30}, {"max_grid": 8}, {"max_grid": 64},
    {"fuse_first": 1}, {"fuse_prolong": 0}, {"fuse_p": 1},
    {"fuse_first": 1, "fuse_prolong": 1, "fuse_p": 1}, {"coarse_max": 1024},
    {"compact": 0}, {"compact": 0, "fuse_first": 1, "fuse_prolong": 1, "fuse_p": 1},
    {"coarse_lds": 0}, {"coarse_lds": 0, "compact": 0, "fuse_prolong": 1},
    {"restrict_rows": 1}, {"restrict_rows": 1, "compact": 0, "max_grid": 8},
    {"restrict_rows": 0}, {"restrict_rows": 0, "compact": 0, "fuse_prolong": 0},
    {"fuse_first": 2}, {"restrict_coarse": 0}, {"restrict_coarse": 1},
    {"fuse_first": 2, "restrict_coarse": 0, "compact": 0},
], ids=lambda k: "-".join(f"{a}{b}" for a, b in k.items()))
@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_schedule_knobs_agree(dev, knobs, precision):
    """Where the coarse levels run, how many workgroups stride over the
    virtual blocks and which stages are fused must not change a bit of x."""
    if _emu(dev) and precision == "mixed" and len(knobs) < 3:
        pytest.skip("on the emulator the all-fused case covers MIXED")
    if _emu(dev) and knobs in EMU_COVERED:
        pytest.skip("on the emulator a combined setting covers this knob")
    sysd = SYSTEMS["c2_3d_small"](_emu(dev))
    xa, outa, _ = gpu_solve(sysd, dev, precision)
    try:
        set_knobs(**knobs)
        xb, outb, _ = gpu_solve(sysd, dev, precision)
    finally:
        set_knobs(**KNOB_DEFAULTS)
    assert outa["iters"] == outb["iters"]
    assert np.array_equal(xa, xb)


def test_mixed_error_vs_f64(dev):
    sysd = scenes.c2_3d(seed=1, n=16, L=3) if _emu(dev) else SYSTEMS["c2_3d_small"](False)
    x, out, _ = gpu_solve(sysd, dev, "mixed")
    x64, out64, _ = gpu_solve(sysd, dev, "f64", rel_tol=1e-24)
    xo, outo, _ = osys.solve(sysd, core.F64, max_iters=500, rel_tol=1e-24)
    assert out64["status"] == 1
    assert rel_l2(x64, xo) < 1e-10
    assert out["true_rr"] < out["tol"]
    assert rel_l2(x, x64) <= 1e-5


def test_max_iters_and_zero_rhs(dev):
    sysd = scenes.c2_3d(seed=1, n=32, L=3)
    x, out, _ = gpu_solve(sysd, dev, "plain", max_iters=1)
    xr, outr, _ = osys.solve(sysd, core.PLAIN, max_iters=1)
    assert out["status"] == 2 and out["iters"] == 1 == outr["iters"]
    assert rel_l2(x, xr) <= 1e-6
    x0, out0, _ = gpu_solve(sysd, dev, "plain", max_iters=0)
    assert out0["iters"] == 0 and out0["status"] == 2 and not np.any(x0)
    z = dict(sysd)
    z["velocity"] = [np.zeros_like(v) for v in sysd["velocity"]]
    z["p0"] = 0.0
    xz, outz, _ = gpu_solve(z, dev, "plain")
    assert outz["status"] == 1 and outz["iters"] == 0 and not np.any(xz)


def _assembled(sysd, dev, precision):
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    d = systems.upload(sysd, dev)
    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], sysd["nu"], sysd["bottom"],
                                 precision=precision, device=dev)
    systems.assemble(s, d)
    return s


@pytest.mark.parametrize("kind", ["pressure", "volume", "psi", "surface"])
def test_level_coupling_detection(dev, kind):
    """Pressure-type operators couple every FLUID pair with -scale_A at every
    level (pressure_project.py:55-57), so every level is detected uniform
    with that value; the potential and surface-tension couplings vary."""
    emu = _emu(dev)
    res = [16, 16, 32] if emu else [32, 48, 64]
    sysd = scenes.pool_sphere_system(res, 7, L=3, kind=kind)
    s = _assembled(sysd, dev, "plain")
    s.load()
    for l in range(sysd["L"]):
        canon = s.Ax[l].cpu().numpy()
        vals = np.unique(canon[canon.view(np.uint32) != 0])
        uni, coef = s.level_coupling(l)
        assert uni == (len(vals) <= 1), (l, vals[:4])
        if kind in ("pressure", "volume"):
            assert uni and coef == np.float32(-sysd["scale_A"])
        elif l == 0:
            assert not uni


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_batch_uneven_equals_alone(dev, precision):
    """A batch mixing the four system kinds at uneven sizes: every result is
    bit-identical to the same system solved alone, and matches the oracle."""
    from mgpcg_amd import batch

    emu = _emu(dev)
    count, div, L = (3, 4, 3) if emu else (8, 2, 4)
    mix = scenes.c4_batch(first_seed=100, count=count, L=L)
    for sysd in mix:  # smaller sides (still divisible by 2^(L-1))
        sysd.update(scenes.pool_sphere_system([n // div for n in sysd["res"]], sysd["seed"], L=L,
                                              kind=sysd["kind"]))
    solvers = [_assembled(sysd, dev, precision) for sysd in mix]
    res = batch.solve_batch(solvers, max_iters=500, rel_tol=1e-12)
    xs = [_x(s) for s in solvers]
    for sysd, xb, rb in zip(mix, xs, res):
        print(f"case {sysd['name']} {precision}", flush=True)
        xa, ra, _ = gpu_solve(sysd, dev, precision)
        assert ra["iters"] == rb["iters"] and ra["status"] == rb["status"]
        assert np.array_equal(xa, xb)
        xr, rr, _ = osys.solve(sysd, PREC[precision], max_iters=500, rel_tol=1e-12)
        assert abs(rb["iters"] - rr["iters"]) <= 1 and rb["status"] == rr["status"]
        assert rel_l2(xb, xr) <= 1e-5


@pytest.mark.parametrize("depth,streams", [(2, 2), (1, 2), (2, 1), (3, 3)],
                         ids=lambda v: str(v))
def test_server_continuous_batching_equals_alone(dev, depth, streams):
    """Requests joining at different iterations (continuous batching) give the
    same bits as solving alone, however many iterations the server keeps
    queued ahead of the host and however many streams it spreads its members
    over; slots and hierarchies are reused."""
    from mgpcg_amd import serve
    from mgpcg_amd import systems

    if _emu(dev) and depth == 3:
        pytest.skip("the emulator covers depths 1 and 2")
    set_knobs(server_depth=depth, server_streams=streams)

    emu = _emu(dev)
    nreq = 3 if emu else 6
    lo, hi = (16, 32) if emu else (32, 64)
    reqs = [scenes.c5_request(1000 + i, lo=lo, hi=hi, L=3) for i in range(nreq)]
    srv = serve.Server(hbm_budget_bytes=4 << 30, precision="mixed", n_levels=3,
                       max_active=3, device=dev)
    devs = [systems.upload(r, dev) for r in reqs]
    done = []
    i = 0
    set_knobs(probe_every=1)
    srv.probe(True)
    while len(done) < len(reqs):
        if i < len(reqs):  # one new request per step: staggered admission
            srv.submit(devs[i], meta=i)
            i += 1
        for rq in srv.step():
            rq["x"] = _x(rq["solver"])
            rq["coarse_from"] = rq["solver"].info(0)["coarse_from"]
            srv.retire(rq)
            done.append(rq)
    n_launch, ms, cells, fluid = srv.probe(False, fluid=True)
    set_knobs(probe_every=4, server_depth=2, server_streams=2)
    assert srv.max_members >= 2  # they really overlapped
    # with every launch timed, the probe counts exactly the level-0 cells swept
    # by smooth(0, 0): two launches per V-cycle (nu = 2), iters + 1 V-cycles
    # per solve, none for members that finished on the device but were not
    # yet retired, and none for systems run entirely by the coarse kernel;
    # likewise their FLUID cells (counted by the packing kernel)
    swept = [rq for rq in done if rq["coarse_from"] > 0]
    expect = sum((rq["result"]["iters"] + 1) * 2 * int(np.prod(reqs[rq["meta"]]["res"]))
                 for rq in swept)
    expect_fluid = sum((rq["result"]["iters"] + 1) * 2 *
                       int(np.count_nonzero(reqs[rq["meta"]]["cell_type"] == 0)) for rq in swept)
    assert cells == expect
    assert fluid == expect_fluid
    assert n_launch > 0 and ms >= 0.0
    for rq in done:
        sysd = reqs[rq["meta"]]
        xa, ra, _ = gpu_solve(sysd, dev, "mixed")
        assert ra["iters"] == rq["result"]["iters"]
        assert np.array_equal(xa, rq["x"])


@pytest.mark.parametrize("depth", [1, 2])
def test_server_table_growth_keeps_queued_readbacks(dev, depth):
    """A lane whose member tables grow (a ninth member joins) while iterations
    are still queued keeps those iterations' readbacks: growing reallocates
    the readback slots, and the slots of queued iterations must carry their
    states over, else the host reads whatever the new allocation held (the CPU
    emulation fills new allocations with a nonzero pattern, which reads as a
    finished system).  Every result must equal solving alone."""
    from mgpcg_amd import serve
    from mgpcg_amd import systems

    set_knobs(server_depth=depth, server_streams=1)
    try:
        emu = _emu(dev)
        lo, hi = (16, 24) if emu else (32, 64)
        reqs = [scenes.c5_request(2000 + i, lo=lo, hi=hi, L=3) for i in range(10)]
        devs = [systems.upload(r, dev) for r in reqs]
        srv = serve.Server(hbm_budget_bytes=4 << 30, precision="mixed", n_levels=3, device=dev)
        done = []

        def step():
            for rq in srv.step():
                rq["x"] = _x(rq["solver"])
                srv.retire(rq)
                done.append(rq)

        for i in range(3):
            srv.submit(devs[i], meta=i)
        step()
        step()
        assert len(srv.running) == 3 and not done  # running, iterations queued
        for i in range(3, len(reqs)):
            srv.submit(devs[i], meta=i)
        steps = 0
        while srv.busy():
            steps += 1
            assert steps < 2000
            step()
        assert srv.max_members >= 9  # the lane's tables grew while it ran
        assert len(done) == len(reqs)
        for rq in done:
            sysd = reqs[rq["meta"]]
            print(f"case {rq['meta']} {sysd['name']}", flush=True)
            xa, ra, _ = gpu_solve(sysd, dev, "mixed")
            assert rq["result"]["status"] == ra["status"], rq["result"]
            assert ra["iters"] == rq["result"]["iters"]
            assert np.array_equal(xa, rq["x"])
    finally:
        set_knobs(server_depth=2, server_streams=2)


@pytest.fixture(params=[1, 2], ids=["newest_first", "round_robin"])
def deferred_streams(dev, request):
    """On the CPU emulation: the library's streams queue their work and run it
    only when the host waits, either newest stream first or one operation of
    each stream in turn (legal orders, as far from issue order as the events
    allow), so a missing event between streams changes the results.  On the
    GPU: the streams as they are (the second parameter repeats the test)."""
    from mgpcg_amd import _lib

    lib = _lib.load()
    if _emu(dev):
        lib.hipemu_set_deferred(request.param)
    yield
    if _emu(dev):
        lib.hipemu_set_deferred(0)


def test_server_resubmit_on_other_lane_equals_alone(dev, deferred_streams):
    """A system that retires from one lane while that lane still has
    iterations queued (server depth 2) and is resubmitted at once lands on the
    other lane; its second solve must still equal solving it alone.  The
    queued iterations of the first lane read the system's state, so the second
    lane's reset of that state must wait for them (the retirement event)."""
    from mgpcg_amd import serve
    from mgpcg_amd import systems

    set_knobs(server_depth=2, server_streams=2)
    try:
        emu = _emu(dev)
        f = 1 if emu else 2
        # A: 7 iterations, B: 1, C: ~36 (oracle, MIXED); shapes differ, so the
        # server's idle pool hands A's own hierarchy back for its resubmission
        a = scenes.pool_sphere_system([16 * f, 16 * f, 16 * f], 1, L=3, kind="pressure")
        b = scenes.pool_sphere_system([16 * f, 16 * f, 32 * f], 1, L=3, kind="surface")
        c = scenes.pool_sphere_system([16 * f, 32 * f, 16 * f], 1, L=3, kind="psi")
        da, db, dc = (systems.upload(x, dev) for x in (a, b, c))
        srv = serve.Server(hbm_budget_bytes=4 << 30, precision="mixed", n_levels=3, device=dev)
        for d, m in ((da, "a"), (db, "b"), (dc, "c")):  # lanes: a -> 0, b -> 1, c -> 0
            srv.submit(d, meta=m)
        first_a = None
        results = {}
        lanes_checked = False
        steps = 0
        while srv.busy():
            steps += 1
            assert steps < 2000
            for rq in srv.step():
                m = rq["meta"]
                rq["x"] = _x(rq["solver"])
                if m == "a" and first_a is None:
                    first_a = rq["solver"]
                    running = sorted(r["meta"] for _, r in srv.running.values())
                    # c still runs on lane 0 and lane 1 is empty: the
                    # resubmission goes to lane 1
                    lanes_checked = running == ["c"] and "b" in results
                    srv.retire(rq)
                    srv.submit(da, meta="a2")
                    continue
                results[m] = rq
                srv.retire(rq)
        assert lanes_checked, "the resubmission did not take the other lane"
        assert results["a2"]["solver"] is first_a if "solver" in results["a2"] else True
        for m, sysd in (("a2", a), ("b", b), ("c", c)):
            print(f"case {m} {sysd['name']}", flush=True)
            xa, ra, _ = gpu_solve(sysd, dev, "mixed")
            assert ra["iters"] == results[m]["result"]["iters"], m
            assert np.array_equal(xa, results[m]["x"]), m
    finally:
        set_knobs(server_depth=2, server_streams=2)
