"""Cell classification and fast-sweeping redistancing (fluid_simulator.py:151-159,
337-341; level_set.py:260-333): the library against the oracle bit for bit,
and the oracle against closed forms.

Closed forms: a level set that is linear along one axis redistances to the
exact signed distance to its plane (target_surface's estimate is exact for a
linear phi and every one-dimensional Eikonal update adds exactly dx); a
sphere redistances to within one cell of the exact distance once the sweeps
have converged, keeps every sign, and the cells next to the interface keep
their target_surface estimates.
"""
import numpy as np
import pytest
import torch

from oracle import levelset as OL

FLUID, AIR, SOLID = 0, 1, 2


def _t(a, dev):
    # a copy: the device calls update in place, and on the CPU build .to() would share memory
    return torch.from_numpy(np.array(a, copy=True)).to(dev)


def _sphere(res, c, r, dx):
    I = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    return np.sqrt(sum(((I[a] + 0.5) * dx - c[a]) ** 2 for a in range(len(res)))) - r


# -- oracle against closed forms --------------------------------------------------
@pytest.mark.parametrize("res,axis", [((20, 16), 0), ((20, 16), 1), ((12, 14, 10), 2),
                                      ((12, 14, 10), 0)])
def test_oracle_plane_is_exact(res, axis):
    dx = 1.0 / 16
    x = (np.arange(res[axis]) + 0.5) * dx
    shape = [1] * len(res)
    shape[axis] = res[axis]
    d = np.broadcast_to((x - 0.41).reshape(shape), res)
    out = OL.redistance((3.0 * d).astype(np.float32), dx)
    assert np.max(np.abs(out - d)) < 1e-5


def test_oracle_sphere_within_a_cell():
    n = 24
    dx = 1.0 / n
    d = _sphere((n, n, n), (0.5, 0.45, 0.55), 0.27, dx)
    phi = (2.5 * d).astype(np.float32)
    out = OL.redistance(phi, dx, repeat_times=4)
    assert np.all(np.sign(out) == np.sign(d))
    assert np.max(np.abs(out - d)) < 1.0 * dx
    # the reference's two repeats have converged next to the interface
    out2 = OL.redistance(phi, dx, repeat_times=2)
    near = np.abs(d) < 3 * dx
    assert np.max(np.abs(out2 - d)[near]) < 0.5 * dx
    tmp, fixed = OL.target_surface(phi, dx)
    assert fixed.any() and np.array_equal(out2[fixed], tmp[fixed])


def test_oracle_markers_and_boundary():
    rng = np.random.default_rng(0)
    ct = rng.integers(0, 3, (9, 7)).astype(np.int8)
    phi = rng.uniform(-1, 1, (9, 7)).astype(np.float32)
    out = OL.apply_markers(ct, phi)
    assert np.all(out[ct == SOLID] == SOLID)
    free = ct != SOLID
    assert np.all(out[free & (phi <= 0)] == FLUID) and np.all(out[free & (phi > 0)] == AIR)
    b = OL.init_boundary(np.zeros((5, 6, 7), np.int8))
    assert b.sum() == 2 * (5 * 6 * 7 - 3 * 4 * 5) and b[1:-1, 1:-1, 1:-1].sum() == 0


# -- library against the oracle ---------------------------------------------------
@pytest.mark.parametrize("res", [(24, 20), (14, 12, 16)], ids=["2d", "3d"])
def test_markers_and_boundary_match_oracle(dev, res):
    from mgpcg_amd import cells

    rng = np.random.default_rng(1)
    ct = rng.integers(0, 3, res).astype(np.int8)
    phi = rng.uniform(-1, 1, res).astype(np.float32)
    phi[rng.uniform(size=res) < 0.05] = 0.0
    t = _t(ct, dev)
    cells.init_boundary(t)
    cells.apply_markers(t, _t(phi, dev))
    ref = OL.apply_markers(OL.init_boundary(ct), phi)
    assert np.array_equal(t.cpu().numpy(), ref)


@pytest.mark.parametrize("res", [(24, 20), (14, 12, 16), (16, 16, 16)],
                         ids=["2d", "3d", "3d_cube"])
@pytest.mark.parametrize("repeat", [2, 3])
def test_redistance_matches_oracle(dev, res, repeat):
    from mgpcg_amd.level_set import FastSweepingLevelSet

    dx = 1.0 / max(res)
    c = [0.5 * n * dx for n in res]
    d = np.minimum(_sphere(res, c, 0.3, dx),
                   (np.meshgrid(*[np.arange(n) for n in res], indexing="ij")[1] + 0.5) * dx
                   - 0.2)
    rng = np.random.default_rng(2)
    phi = (d * rng.uniform(0.5, 3.0, res)).astype(np.float32)  # not a distance
    phi[rng.uniform(size=res) < 0.01] = 0.0
    ls = FastSweepingLevelSet(len(res), res, dx, phi=_t(phi, dev), device=dev)
    ls.repeat_times = repeat
    ls.redistance()
    ref = OL.redistance(phi, dx, repeat)
    assert np.array_equal(ls.phi.cpu().numpy(), ref)
