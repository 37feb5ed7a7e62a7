"""apply_pressure / apply_magnetic_force (fluid_simulator.py:192-214,
ferrofluid_simulator.py:114-137): the library against the oracle bit for
bit, and the projection identity that pins both against the mathematics.

Identity: with p0 = 0, re-assembling the right-hand side from the projected
velocity gives b(u - scale grad p) = b(u) - A p = the solve's residual (the
update of each FLUID-FLUID face is sA (p_j - p_i) / sb, of each FLUID-AIR
face the ghost-fluid weighted sA c (p0 - p_i) / sb, and SOLID faces do not
enter b).  So after a converged solve the projected field is discretely
divergence-free to the solve tolerance; a wrong sign, index or weight in
either kernel leaves an O(1) divergence.
"""
import numpy as np
import pytest
import torch

from mgpcg_amd import scenes
from helpers import rel_l2
from oracle import strategies as S

FLUID, AIR, SOLID = 0, 1, 2


def _fields(res, seed, dev):
    rng = np.random.default_rng(seed)
    dim = len(res)
    u = rng.uniform(size=res)
    ct = np.where(u < 0.5, FLUID, np.where(u < 0.8, AIR, SOLID)).astype(np.int8)
    phi = rng.uniform(-1.0, 1.0, res).astype(np.float32)
    phi[rng.uniform(size=res) < 0.02] = 0.0  # phi_i == 0 takes the clamp value
    p = rng.standard_normal(res).astype(np.float32)
    vel = [rng.uniform(-1, 1, [n + (a == k) for a, n in enumerate(res)]).astype(np.float32)
           for k in range(dim)]
    H = [rng.uniform(-2, 2, [n + (a == k) for a, n in enumerate(res)]).astype(np.float32)
         for k in range(dim)]
    return ct, phi, p, vel, H


def _t(a, dev):
    # a copy: the device calls update in place, and on the CPU build .to() would share memory
    return torch.from_numpy(np.array(a, copy=True)).to(dev)


@pytest.mark.parametrize("res", [(24, 20), (12, 10, 14)], ids=["2d", "3d"])
@pytest.mark.parametrize("gfm", [False, True], ids=["plain", "gfm"])
def test_apply_pressure_matches_oracle(dev, res, gfm):
    from mgpcg_amd import projection

    ct, phi, p, vel, _ = _fields(res, 3, dev)
    dt, rho, dx, p0, c_max = 1.25e-2, 1000.0, 1.0 / 32, 0.7, 50.0
    ref = S.apply_pressure(ct, vel, p, dt, rho, dx, p0, gfm, phi, c_max)
    v = [_t(a, dev) for a in vel]
    projection.apply_pressure(v, _t(ct, dev), _t(p, dev), dt, rho, dx, p0, gfm,
                              _t(phi, dev) if gfm else None, c_max)
    for k in range(len(res)):
        assert np.array_equal(v[k].cpu().numpy(), ref[k]), k
    # the top faces u_k[res_k] are never written (the reference loops over cells)
    for k in range(len(res)):
        top = [slice(None)] * len(res)
        top[k] = -1
        assert np.array_equal(v[k].cpu().numpy()[tuple(top)], vel[k][tuple(top)])


@pytest.mark.parametrize("res", [(24, 20), (12, 10, 14)], ids=["2d", "3d"])
def test_apply_magnetic_force_matches_oracle(dev, res):
    from mgpcg_amd import projection

    ct, _, p, vel, H = _fields(res, 4, dev)
    dt, rho, dx, p0, mu0, kk = 1.25e-2, 1000.0, 1.0 / 32, 0.3, 1.25663706e-6, 2.0
    ref = S.apply_magnetic_force(ct, vel, p, H, dt, rho, dx, p0, mu0, kk)
    v = [_t(a, dev) for a in vel]
    projection.apply_magnetic_force(v, _t(ct, dev), _t(p, dev), [_t(h, dev) for h in H], dt,
                                    rho, dx, p0, mu0, kk)
    for k in range(len(res)):
        assert np.array_equal(v[k].cpu().numpy(), ref[k]), k


def _pressure_b(sysd, dev, velocity, ct, phi):
    from mgpcg_amd import _lib

    lib = _lib.load()
    b = torch.empty(sysd["res"], dtype=torch.float32, device=dev)
    _lib.check(lib.mgpcg_pressure_build_b(
        _lib.stream_handle(dev), sysd["dim"], _lib.int3(sysd["res"]), _lib.ptr(ct),
        _lib.ptr_array(velocity), _lib.ptr(phi), float(sysd["scale_A"]),
        float(sysd["scale_b"]), float(sysd["p0"]), int(sysd["gfm"]), float(sysd["c_max"]),
        _lib.ptr(b)))
    return b.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_projection_leaves_residual_divergence(dev, precision):
    from mgpcg_amd import mgpcg
    from mgpcg_amd import projection
    from mgpcg_amd import systems

    emu = dev == "cpu"
    sysd = scenes.c2_3d(seed=5, n=24 if emu else 64, L=3 if emu else 4)
    sysd["p0"] = 0.0  # the identity needs p0 = 0
    d = systems.upload(sysd, dev)
    s = mgpcg.MGPCGPoissonSolver(3, sysd["res"], sysd["L"], 2, 10, precision=precision,
                                 device=dev)
    systems.assemble(s, d)
    b0 = s.b.cpu().numpy().astype(np.float64)
    s.solve(500, rel_tol=1e-12)
    rel_res = (s.last["rr"] / s.last["r0r0"]) ** 0.5
    vel = [u.clone() for u in d["velocity"]]
    ct = d["cell_type"].to(torch.int8)
    projection.apply_pressure(vel, ct, s.x, sysd["dt"], sysd["rho"], sysd["dx"], 0.0, True,
                              d["phi"], sysd["c_max"])
    b1 = _pressure_b(sysd, dev, vel, ct, d["phi"])
    fluid = sysd["cell_type"] == FLUID
    before = np.linalg.norm(b0[fluid])
    after = np.linalg.norm(b1[fluid])
    print(f"|b(u)| {before:.3e} -> |b(u_new)| {after:.3e}, solve residual {rel_res:.2e}",
          flush=True)
    # |b(u_new)| is the solve's residual norm, up to FP32 rounding of the update
    assert after <= (10.0 * rel_res + 1e-6) * before
    # without the gradient step the identity fails by O(1): the check has teeth
    b_same = _pressure_b(sysd, dev, [u.clone() for u in d["velocity"]], ct, d["phi"])
    assert rel_l2(b_same[fluid], b0[fluid]) < 1e-6


# -- extrap_velocity (fluid_simulator.py:240-270) ---------------------------------
@pytest.mark.parametrize("res", [(20, 18), (12, 10, 14)], ids=["2d", "3d"])
@pytest.mark.parametrize("iters", [1, 10])
def test_extrap_velocity_matches_oracle(dev, res, iters):
    from mgpcg_amd import projection
    from oracle import levelset as OL

    rng = np.random.default_rng(7)
    ct, _, _, vel, _ = _fields(res, 8, dev)
    ct = np.where(rng.uniform(size=res) < 0.7, AIR, ct).astype(np.int8)  # sparse fluid
    ref = OL.extrap_velocity(ct, vel, iters)
    v = [_t(a, dev) for a in vel]
    projection.extrap_velocity(v, _t(ct, dev), iters)
    for k in range(len(res)):
        assert np.array_equal(v[k].cpu().numpy(), ref[k]), k


def test_extrap_velocity_closed_form():
    """A fluid blob with one constant velocity extends that constant to every
    face within reach and leaves the valid faces unchanged; faces beyond the
    iteration count keep their values."""
    from oracle import levelset as OL

    res = (24, 24)
    ct = np.full(res, AIR, np.int8)
    ct[10:14, 10:14] = FLUID
    vel = [np.full([n + (a == k) for a, n in enumerate(res)], 9.0, np.float32) for k in range(2)]
    for k in range(2):
        sl = [slice(10, 14), slice(10, 14)]
        sl[k] = slice(10, 15)
        vel[k][tuple(sl)] = 0.25 * (k + 1)
    out = OL.extrap_velocity(ct, vel, 3)
    for k in range(2):
        c = 0.25 * (k + 1)
        sl = [slice(10, 14), slice(10, 14)]
        sl[k] = slice(10, 15)
        assert np.all(out[k][tuple(sl)] == np.float32(c))  # valid faces unchanged
        reach = [slice(7, 17), slice(7, 17)]
        reach[k] = slice(7, 18)
        assert np.all(out[k][tuple(reach)] == np.float32(c))  # three layers out
        assert np.all(out[k][:5, :] == 9.0) and np.all(out[k][:, :5] == 9.0)  # beyond reach
