"""Pressure systems assembled straight into the packed hierarchy
(MGPCGPoissonSolver.reinitialize_packed over mgpcg_assemble_pressure) against
the canonical route (reinitialize: the strategy's build_b / build_A kernels
write the canonical fields, then load packs them).

The two routes must give the same device hierarchy bit for bit: every level's
flag words, FLUID block flags, work orders, coupling summary and coarse cell
types, and then the same solve (iterations, residual norms, x).  Covered: the
ghost-fluid pressure system (pressure_project.py:45-64, 15-39), the plain one
(no ghost fluid, 2D), the volume-controlled one with a given divergence
constant and with the PI controller choosing it (volume_control.py:39-51),
ragged shapes, and each precision.  Runs on the MI355X and on the CPU
emulation build.
"""
import numpy as np
import pytest
import torch


def _pair(dev, sysd, precision, L=None):
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    L = L or sysd["L"]
    d = systems.upload(sysd, dev)
    a = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], L, 2, 10, precision=precision,
                                 device=dev)
    b = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], L, 2, 10, precision=precision,
                                 device=dev)
    systems.assemble(a, d)
    a.load()
    systems.assemble(b, d, packed=True)
    b.load()  # nothing to do on the packed route
    return a, b


def _same_hierarchy(a, b):
    for l in range(a.n_mg_levels):
        assert torch.equal(a.grid_type[l].cpu(), b.grid_type[l].cpu()), l
        fa, wa = a.level_flags(l)
        fb, wb = b.level_flags(l)
        assert np.array_equal(fa, fb), l
        assert np.array_equal(wa, wb), l
        oa, ob = a.level_order(l), b.level_order(l)
        assert np.array_equal(oa[0], ob[0]), l
        assert (oa[1] is None) == (ob[1] is None), l
        if oa[1] is not None:
            assert np.array_equal(oa[1], ob[1]), l
        ua, ca = a.level_coupling(l)
        ub, cb = b.level_coupling(l)
        assert ua == ub and np.float32(ca).view(np.uint32) == np.float32(cb).view(np.uint32), l


def _same_solve(a, b, iters=60):
    a.solve(iters, rel_tol=1e-12)
    b.solve(iters, rel_tol=1e-12)
    for k in ("iters", "status", "r0r0", "rr", "true_rr"):
        assert a.last[k] == b.last[k], (k, a.last[k], b.last[k])
    assert torch.equal(a.x.cpu(), b.x.cpu())
    if a.x64 is not None:
        assert torch.equal(a.x64.cpu(), b.x64.cpu())
    assert a.last["r0r0"] > 0


@pytest.mark.parametrize("precision", ["plain", "mixed", "f64"])
def test_ghost_fluid_pressure(dev, precision):
    from mgpcg_amd import scenes

    sysd = scenes.c2_3d(seed=3, n=24, L=3)
    a, b = _pair(dev, sysd, precision)
    _same_hierarchy(a, b)
    _same_solve(a, b)


def test_plain_pressure_2d(dev):
    from mgpcg_amd import scenes

    sysd = scenes.c1_2d(seed=4, n=48)
    assert not sysd["gfm"]
    a, b = _pair(dev, sysd, "mixed", L=4)
    _same_hierarchy(a, b)
    _same_solve(a, b)


@pytest.mark.parametrize("res", [[20, 24, 16], [20, 28, 36], [36, 28]])
def test_ragged_shapes(dev, res):
    """Odd sides at the coarse levels and padded rows of every width."""
    from mgpcg_amd import scenes

    sysd = scenes.pool_sphere_system(res, seed=7, L=3)
    a, b = _pair(dev, sysd, "mixed")
    _same_hierarchy(a, b)
    _same_solve(a, b)


@pytest.mark.parametrize("res,L", [([32, 32, 32], 6), ([64, 96], 6), ([128, 64], 7)])
def test_deep_hierarchies(dev, res, L):
    """Hierarchies of 6 and 7 levels, 3D and 2D: one packing launch over all
    of them."""
    from mgpcg_amd import scenes

    sysd = scenes.pool_sphere_system(res, seed=13, L=L)
    a, b = _pair(dev, sysd, "mixed")
    _same_hierarchy(a, b)
    _same_solve(a, b)


def test_seeded_volume_control(dev):
    from mgpcg_amd import scenes

    sysd = scenes.pool_sphere_system([24, 24, 24], seed=9, L=3, kind="volume")
    assert sysd["c"] != 0.0
    a, b = _pair(dev, sysd, "mixed")
    _same_hierarchy(a, b)
    _same_solve(a, b)


def test_pi_controller_on_both_routes(dev):
    """The PI controller runs once per assembly on either route: the first
    assembly records the target volume (no constant added), the second adds
    the constant the changed level set asks for, and both routes agree."""
    from mgpcg_amd import mgpcg
    from mgpcg_amd import scenes
    from mgpcg_amd import systems
    from mgpcg_amd.level_set import LevelSet
    from mgpcg_amd.volume_control import PressureProjectWithVolumeControlStrategy

    sysd = scenes.c2_3d(seed=5, n=24, L=3)
    d = systems.upload(sysd, dev)
    res = sysd["res"]
    routes = []
    for packed in (False, True):
        s = mgpcg.MGPCGPoissonSolver(3, res, 3, 2, 10, precision="mixed", device=dev)
        phi = d["phi"].clone()
        ls = LevelSet(3, res, sysd["dx"], phi=phi, device=s.device)
        st = PressureProjectWithVolumeControlStrategy(3, d["velocity"], True, phi, sysd["p0"],
                                                      ls, sysd["dt"], c_max=sysd["c_max"])
        st.scale_A, st.scale_b = sysd["scale_A"], sysd["scale_b"]
        cs = []
        for step in range(2):
            if step == 1:
                phi.add_(0.25 * sysd["dx"])  # the fluid lost volume
            if packed:
                s.reinitialize_packed(d["cell_type"], st)
            else:
                s.reinitialize(d["cell_type"], st)
            cs.append(st.last_c)
            s.load()
            s.solve(60, rel_tol=1e-12)
        routes.append((s, cs))
    (a, ca), (b, cb) = routes
    assert ca == cb and ca[0] is None and ca[1] is not None and ca[1] != 0.0
    _same_hierarchy(a, b)
    for k in ("iters", "status", "r0r0", "rr"):
        assert a.last[k] == b.last[k], k
    assert torch.equal(a.x.cpu(), b.x.cpu())


def test_route_switches(dev):
    """A packed solve can be repeated (the hierarchy keeps b), and a later
    canonical reinitialize takes the canonical fields again."""
    from mgpcg_amd import scenes
    from mgpcg_amd import systems

    s1 = scenes.pool_sphere_system([16, 24, 16], seed=11, L=3)
    s2 = scenes.pool_sphere_system([16, 24, 16], seed=12, L=3)
    a, b = _pair(dev, s1, "mixed")
    _same_solve(a, b)
    x1 = b.x.cpu().clone()
    b.solve(60, rel_tol=1e-12)
    assert torch.equal(b.x.cpu(), x1)
    d2 = systems.upload(s2, dev)
    systems.assemble(a, d2, packed=True)
    systems.assemble(b, d2)
    b.load()
    _same_hierarchy(a, b)
    _same_solve(a, b)


def test_assemble_pressure_rejects_bad_arguments(dev):
    from mgpcg_amd import _lib
    from mgpcg_amd import mgpcg
    from mgpcg_amd import scenes
    from mgpcg_amd import systems

    sysd = scenes.c2_3d(seed=3, n=16, L=3)
    d = systems.upload(sysd, dev)
    s = mgpcg.MGPCGPoissonSolver(3, sysd["res"], 3, 2, 10, precision="mixed", device=dev)
    lib = _lib.load()
    with pytest.raises(RuntimeError, match="ghost-fluid method needs phi"):
        _lib.check(lib.mgpcg_assemble_pressure(
            s._h, _lib.ptr_array(s.grid_type), _lib.ptr_array(d["velocity"]), None, 1.0, 1.0,
            0.0, 1, 100.0, 0, 0.0))
    with pytest.raises(RuntimeError, match="velocity"):
        _lib.check(lib.mgpcg_assemble_pressure(
            s._h, _lib.ptr_array(s.grid_type), _lib.ptr_array([d["velocity"][0], None, None]),
            _lib.ptr(d["phi"]), 1.0, 1.0, 0.0, 1, 100.0, 0, 0.0))
