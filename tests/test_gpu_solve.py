"""Whole-solve parity of the HIP solver against the oracle (GPU).

Bar: pressures within 1e-5 relative L2 of the oracle at 1e-6 relative
residual (rel_tol 1e-12 on squared norms, mgpcg.py:164), iteration count
within +-1, same status; MIXED additionally within 1e-5 of an FP64 solve
converged to 1e-12 relative residual.
"""
import numpy as np
import pytest
import torch

import scenes
from helpers import rel_l2
from oracle import core
from oracle import systems as osys

pytestmark = pytest.mark.gpu

PREC = {"plain": core.PLAIN, "mixed": core.MIXED, "f64": core.F64}


def gpu_solve(sysd, precision="plain", hostsync=False, max_iters=500, rel_tol=1e-12, opts=None):
    import mgpcg
    import systems

    dev = systems.upload(sysd, "cuda:0")
    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], sysd["nu"], sysd["bottom"],
                                 precision=precision, device="cuda:0")
    for k, v in (opts or {}).items():
        s.set_option(k, v)
    systems.assemble(s, dev)
    if hostsync:
        s.solve_hostsync(max_iters, rel_tol=rel_tol)
    else:
        s.solve(max_iters, rel_tol=rel_tol)
    torch.cuda.synchronize()
    x = (s.x64 if s.x64 is not None else s.x).cpu().numpy().astype(np.float64)
    return x, dict(s.last), s


SYSTEMS = [
    ("c1_2d_128", lambda: scenes.c1_2d(seed=0)),
    ("c2_3d_64", lambda: scenes.c2_3d(seed=1, n=64, L=4)),
    ("c2_3d_128", lambda: scenes.c2_3d(seed=1, n=128, L=5)),
]


@pytest.mark.parametrize("name,make", SYSTEMS, ids=[s[0] for s in SYSTEMS])
@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_solve_matches_oracle(name, make, precision):
    sysd = make()
    print(f"case {name} {precision}", flush=True)
    x, out, _ = gpu_solve(sysd, precision)
    xr, outr, _ = osys.solve(sysd, PREC[precision], max_iters=500, rel_tol=1e-12)
    assert out["status"] == 1 == outr["status"], (out, outr)
    assert abs(out["iters"] - outr["iters"]) <= 1, (out["iters"], outr["iters"])
    assert rel_l2(x, xr) <= 1e-5
    assert out["rr"] < out["tol"]


def test_hostsync_loop_equals_device_loop():
    sysd = scenes.c2_3d(seed=1, n=64, L=4)
    xa, outa, _ = gpu_solve(sysd, "plain")
    xb, outb, _ = gpu_solve(sysd, "plain", hostsync=True)
    assert outa["iters"] == outb["iters"]
    assert np.array_equal(xa, xb)


@pytest.mark.parametrize("use_graph,burst", [(0, 1), (1, 1), (1, 4), (1, 16)])
def test_launch_modes_agree(use_graph, burst):
    import _lib

    sysd = scenes.c2_3d(seed=1, n=64, L=4)
    xa, outa, _ = gpu_solve(sysd, "plain")
    xb, outb, _ = gpu_solve(sysd, "plain",
                            opts={_lib.OPT_USE_GRAPH: use_graph, _lib.OPT_BURST: burst})
    assert outa["iters"] == outb["iters"]
    assert np.array_equal(xa, xb)


def test_mixed_error_vs_f64():
    sysd = scenes.c2_3d(seed=1, n=64, L=4)
    x, out, _ = gpu_solve(sysd, "mixed")
    x64, out64, _ = gpu_solve(sysd, "f64", rel_tol=1e-24)
    xo, outo, _ = osys.solve(sysd, core.F64, max_iters=500, rel_tol=1e-24)
    assert out64["status"] == 1
    assert rel_l2(x64, xo) < 1e-10
    assert out["true_rr"] < out["tol"]
    assert rel_l2(x, x64) <= 1e-5


def test_max_iters_and_zero_rhs():
    sysd = scenes.c2_3d(seed=1, n=32, L=3)
    x, out, _ = gpu_solve(sysd, "plain", max_iters=1)
    xr, outr, _ = osys.solve(sysd, core.PLAIN, max_iters=1)
    assert out["status"] == 2 and out["iters"] == 1 == outr["iters"]
    assert rel_l2(x, xr) <= 1e-6
    x0, out0, _ = gpu_solve(sysd, "plain", max_iters=0)
    assert out0["iters"] == 0 and out0["status"] == 2 and not np.any(x0)
    z = dict(sysd)
    z["velocity"] = [np.zeros_like(v) for v in sysd["velocity"]]
    z["p0"] = 0.0
    xz, outz, _ = gpu_solve(z, "plain")
    assert outz["status"] == 1 and outz["iters"] == 0 and not np.any(xz)
