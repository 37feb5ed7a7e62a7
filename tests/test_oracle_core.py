"""Pins the C oracle (oracle/mgpcg_oracle.c) of mgpcg.py against closed forms,
brute force and invariants -- no reference golden vectors exist (the
reference ships no tests and Taichi is not installable here).

What pins what:
  coarsen         brute force over all 3^(2^d) child patterns with an independent
                  priority formulation (mgpcg.py:46-60)
  compute_Ap      an independently assembled scipy.sparse matrix; the Dirichlet
                  Laplacian's sine eigenvectors (closed form)
  smooth          zero residual on every updated cell; untouched other colour
  v_cycle (L=1)   an independent pointwise red-black Gauss-Seidel on the matrix
  restrict/prol.  the adjoint identity <R u, v> = <u, P v> on FLUID cells
  v_cycle         symmetry and definiteness of the preconditioner
  solve           manufactured solutions and scipy spsolve; loop-count semantics
A planted-mistake test shows the matrix check rejects plausible slips.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from helpers import FLUID, AIR, SOLID, assemble_matrix, random_hierarchy, rel_l2
from oracle import core
from oracle import strategies as S


# ----------------------------------------------------------------------------
# coarsening (init_gridtype)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_coarsen_all_patterns(dim):
    nch = 2 ** dim
    patterns = np.array(list(itertools.product([FLUID, AIR, SOLID], repeat=nch)), np.int8)
    K = len(patterns)
    if dim == 2:
        fine = np.zeros((2 * K, 2), np.int8)
        for p, pat in enumerate(patterns):
            fine[2 * p:2 * p + 2, 0:2] = pat.reshape(2, 2)
    else:
        fine = np.zeros((2, 2 * K, 2), np.int8)
        for p, pat in enumerate(patterns):
            fine[0:2, 2 * p:2 * p + 2, 0:2] = pat.reshape(2, 2, 2)
    coarse = core.coarsen(fine).ravel()
    # independent formulation: AIR outranks FLUID outranks SOLID
    rank = {AIR: 2, FLUID: 1, SOLID: 0}
    unrank = {2: AIR, 1: FLUID, 0: SOLID}
    expect = np.array([unrank[max(rank[int(c)] for c in pat)] for pat in patterns], np.int8)
    assert np.array_equal(coarse, expect)


# ----------------------------------------------------------------------------
# operator
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("res", [(20, 18), (10, 12, 8)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_compute_Ap_matches_sparse_matrix(res, dtype):
    h = random_hierarchy(res, 1, seed=3, dtype=dtype)
    gt, Ad, Ax = h["gt"][0], h["Adiag"][0], h["Ax"][0]
    p = np.random.default_rng(4).standard_normal(res).astype(dtype)
    Ap = core.compute_Ap(gt, Ad, Ax, p)
    A = assemble_matrix(gt, Ad, Ax)
    ref = (A @ p.astype(np.float64).ravel()).reshape(res)
    fl = gt == FLUID
    tol = 2e-6 if dtype == np.float32 else 1e-13
    assert rel_l2(Ap[fl], ref[fl]) < tol
    assert np.all(Ap[~fl] == 0)
    # A is symmetric by construction (one coupling per face)
    assert abs(A - A.T).max() == 0


def _air_ring_box(res):
    ct = np.full(res, FLUID, np.int8)
    for a in range(len(res)):
        sl = [slice(None)] * len(res)
        sl[a] = 0
        ct[tuple(sl)] = AIR
        sl[a] = -1
        ct[tuple(sl)] = AIR
    return ct


@pytest.mark.parametrize("res,modes", [((12, 10), (1, 3)), ((10, 8, 12), (2, 1, 3))])
def test_pressure_operator_dirichlet_eigenvector(res, modes):
    """All-FLUID interior with an AIR ring, GFM off: A = sA * Dirichlet Laplacian,
    so A v = lambda v for v = prod sin(pi k_a i_a / (m_a + 1))."""
    sA = 0.375  # exactly representable in FP32, the strategy's constant type
    ct = _air_ring_box(res)
    Ad, Ax = S.pressure_build_A(ct, 0, sA, ghost_fluid_method=False)
    Ad, Ax = Ad.astype(np.float64), Ax.astype(np.float64)
    grids = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    v = np.ones(res)
    lam = 0.0
    for a, (n, k) in enumerate(zip(res, modes)):
        m = n - 2
        v = v * np.sin(np.pi * k * grids[a] / (m + 1))
        lam += 2 - 2 * np.cos(np.pi * k / (m + 1))
    lam *= sA
    v = np.where(ct == FLUID, v, 0.0)
    Av = core.compute_Ap(ct, Ad, Ax, np.ascontiguousarray(v))
    fl = ct == FLUID
    assert rel_l2(Av[fl], lam * v[fl]) < 1e-12


def _ns_numpy(Ax, x, mistake=None):
    """neighbor_sum written with numpy shifts; `mistake` plants a slip."""
    dim = x.ndim
    ret = np.zeros_like(x)
    for k in range(dim):
        lo = np.zeros_like(x)
        hi = np.zeros_like(x)
        sl_lo = [slice(None)] * dim
        sl_hi = [slice(None)] * dim
        sl_lo[k] = slice(0, x.shape[k] - 1)
        sl_hi[k] = slice(1, x.shape[k])
        if mistake == "own_coeff_below":  # Ax[I][k] instead of Ax[I-e_k][k]
            lo[tuple(sl_hi)] = Ax[..., k][tuple(sl_hi)] * x[tuple(sl_lo)]
        else:
            lo[tuple(sl_hi)] = Ax[..., k][tuple(sl_lo)] * x[tuple(sl_lo)]
        hi[tuple(sl_lo)] = Ax[..., k][tuple(sl_lo)] * x[tuple(sl_hi)]
        if mistake == "drop_upper" and k == dim - 1:
            hi[:] = 0
        if mistake == "sign" and k == 0:
            lo = -lo
        if mistake == "transpose_axes" and dim > 1:
            kk = (k + 1) % dim
            lo[:] = 0
            lo[tuple(sl_hi)] = Ax[..., kk][tuple(sl_lo)] * x[tuple(sl_lo)]
        ret = ret + (lo + hi)
    return ret


@pytest.mark.parametrize("mistake", ["own_coeff_below", "drop_upper", "sign", "transpose_axes"])
def test_matrix_check_rejects_planted_mistakes(mistake):
    res = (9, 8, 7)
    h = random_hierarchy(res, 1, seed=11, dtype=np.float64)
    gt, Ad, Ax = h["gt"][0], h["Adiag"][0], h["Ax"][0]
    p = np.random.default_rng(12).standard_normal(res)
    A = assemble_matrix(gt, Ad, Ax)
    fl = gt == FLUID
    ref = (A @ p.ravel()).reshape(res)
    good = Ad * p + _ns_numpy(Ax, p)
    assert rel_l2(good[fl], ref[fl]) < 1e-13
    bad = Ad * p + _ns_numpy(Ax, p, mistake)
    assert rel_l2(bad[fl], ref[fl]) > 1e-3


# ----------------------------------------------------------------------------
# smoother and V-cycle
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("res", [(16, 14), (8, 10, 6)])
@pytest.mark.parametrize("phase", [0, 1])
def test_smooth_zeroes_residual_on_updated_colour(res, phase):
    h = random_hierarchy(res, 1, seed=5, dtype=np.float64)
    gt, Ad, Ax = h["gt"][0], h["Adiag"][0], h["Ax"][0]
    rng = np.random.default_rng(6)
    r = rng.standard_normal(res)
    z = rng.standard_normal(res)
    z0 = z.copy()
    core.smooth(gt, Ad, Ax, r, z, phase)
    A = assemble_matrix(gt, Ad, Ax)
    resid = r - (A @ z.ravel()).reshape(res)
    colour = np.indices(res).sum(axis=0) & 1
    upd = (gt == FLUID) & (colour == phase)
    assert np.abs(resid[upd]).max() < 1e-12 * max(1.0, np.abs(r).max())
    assert np.array_equal(z[~upd], z0[~upd])


def _pointwise_rbgs(A, gt, r, z, phase_order):
    """Red-black Gauss-Seidel cell by cell from the matrix rows."""
    A = A.tocsr()
    colour = (np.indices(gt.shape).sum(axis=0) & 1).ravel()
    fl = (gt == FLUID).ravel()
    z = z.ravel().copy()
    rv = r.ravel()
    for ph in phase_order:
        new = z.copy()
        for i in np.nonzero(fl & (colour == ph))[0]:
            row = A.getrow(i)
            s = 0.0
            for j, a in zip(row.indices, row.data):
                if j != i:
                    s += a * z[j]
            new[i] = (rv[i] - s) / A[i, i]
        z = new
    return z.reshape(gt.shape)


@pytest.mark.parametrize("bottom", [2, 6])
def test_single_level_vcycle_is_symmetric_gauss_seidel(bottom):
    res = (6, 5, 4)
    h = random_hierarchy(res, 1, seed=7, dtype=np.float64)
    r = np.random.default_rng(8).standard_normal(res)
    z, _, _ = core.v_cycle(h, r, nu=2, bottom=bottom)
    A = assemble_matrix(h["gt"][0], h["Adiag"][0], h["Ax"][0], fluid_only=False)
    order = [0, 1] * (bottom // 2) + [1, 0] * (bottom // 2)
    zr = _pointwise_rbgs(A, h["gt"][0], r, np.zeros(res), order)
    fl = h["gt"][0] == FLUID
    assert rel_l2(z[fl], zr[fl]) < 1e-12


@pytest.mark.parametrize("res", [(16, 12), (8, 8, 12)])
def test_restrict_prolongate_adjoint(res):
    h = random_hierarchy(res, 2, seed=9, dtype=np.float64)
    gt, Ad, Ax = h["gt"][0], h["Adiag"][0], h["Ax"][0]
    rng = np.random.default_rng(10)
    u = rng.standard_normal(res)
    v = rng.standard_normal(tuple(n // 2 for n in res))
    Ru = core.restrict(gt, Ad, Ax, u, np.zeros(res))  # z = 0: plain sum over FLUID children
    Pv = np.zeros(res)
    core.prolongate(Pv, np.ascontiguousarray(v))
    fl = gt == FLUID
    lhs = float(np.sum(Ru * v))
    rhs = float(np.sum(u[fl] * Pv[fl]))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


@pytest.mark.parametrize("res,L", [((32, 24), 3), ((16, 16, 8), 3)])
def test_vcycle_symmetric_positive(res, L):
    h = random_hierarchy(res, L, seed=13, dtype=np.float64)
    fl = h["gt"][0] == FLUID
    rng = np.random.default_rng(14)
    a = np.where(fl, rng.standard_normal(res), 0.0)
    b = np.where(fl, rng.standard_normal(res), 0.0)
    Ma, _, _ = core.v_cycle(h, a)
    Mb, _, _ = core.v_cycle(h, b)
    ab = float(np.sum(Ma[fl] * b[fl]))
    ba = float(np.sum(a[fl] * Mb[fl]))
    assert abs(ab - ba) <= 1e-10 * max(abs(ab), 1.0)
    assert float(np.sum(Ma[fl] * a[fl])) > 0


# ----------------------------------------------------------------------------
# solve
# ----------------------------------------------------------------------------
def _pressure_box(res, seed, gfm=False):
    rng = np.random.default_rng(seed)
    dx = 1.0 / max(res)
    X = np.meshgrid(*[(np.arange(n) + 0.5) * dx for n in res], indexing="ij")
    phi = (X[1] - 0.6 * res[1] * dx).astype(np.float32)
    ct = np.where(phi <= 0, FLUID, AIR).astype(np.int8)
    for a in range(len(res)):
        sl = [slice(None)] * len(res)
        sl[a] = 0
        ct[tuple(sl)] = SOLID
        sl[a] = -1
        ct[tuple(sl)] = SOLID
    sA = 0.8
    L = 3
    h = S.hierarchy_from_cell_type(ct, L, lambda g, l: S.pressure_build_A(g, l, sA, gfm, phi, 1e3))
    return h, rng


@pytest.mark.parametrize("res", [(32, 32), (16, 16, 16)])
def test_solve_manufactured_f64(res):
    h, rng = _pressure_box(res, 21)
    h64 = core.to_f64(h)
    gt = h64["gt"][0]
    fl = gt == FLUID
    xs = np.where(fl, rng.standard_normal(res), 0.0)
    A = assemble_matrix(gt, h64["Adiag"][0], h64["Ax"][0])
    b = (A @ xs.ravel()).reshape(res)
    x, out, _ = core.solve(h64, b, precision=core.F64, max_iters=200, rel_tol=1e-28,
                           abs_tol=1e-300, eps=0.0)
    assert out["status"] == core.PLAIN + 1 or out["status"] == 2
    assert rel_l2(x[fl], xs[fl]) < 1e-9


@pytest.mark.parametrize("res", [(32, 32), (16, 16, 16)])
@pytest.mark.parametrize("precision", [core.PLAIN, core.MIXED])
def test_solve_fp32_meets_tolerance_and_spsolve(res, precision):
    h, rng = _pressure_box(res, 22)
    gt = h["gt"][0]
    fl = gt == FLUID
    b = np.where(fl, rng.uniform(-1, 1, res), 0.0).astype(np.float32)
    x, out, hist = core.solve(h, b, precision=precision, max_iters=500, rel_tol=1e-12,
                              hist_cap=64)
    assert out["status"] == 1
    A = assemble_matrix(gt, h["Adiag"][0], h["Ax"][0])
    idx = np.nonzero(fl.ravel())[0]
    Af = A[idx][:, idx]
    xs = spla.spsolve(Af.tocsc(), b.astype(np.float64).ravel()[idx])
    true_rel = np.linalg.norm(b.ravel()[idx] - Af @ x.ravel()[idx]) / np.linalg.norm(b.ravel()[idx])
    # PLAIN tests the recursively updated FP32 residual (mgpcg.py:188-194), which
    # drifts from the true one; MIXED accepts only on the FP64 true residual.
    assert true_rel < (1e-5 if precision == core.PLAIN else 1e-6)
    assert rel_l2(x.ravel()[idx], xs) < 1e-4
    assert hist[out["iters"]] < out["tol"]
    if precision == core.PLAIN:
        # stops at the first recursive residual under tol; MIXED may go on
        # when the true residual disagrees
        assert np.all(hist[: out["iters"]] >= out["tol"])


def test_solve_loop_count_semantics():
    """iter counts completed non-final loop bodies (mgpcg.py:176-209)."""
    h, rng = _pressure_box((16, 16), 23)
    fl = h["gt"][0] == FLUID
    b = np.where(fl, rng.uniform(-1, 1, (16, 16)), 0.0).astype(np.float32)
    x0, out0, _ = core.solve(h, b, max_iters=0)
    assert out0["iters"] == 0 and out0["status"] == 2 and not np.any(x0)
    x1, out1, _ = core.solve(h, b, max_iters=1)
    assert out1["iters"] == 1 and out1["status"] == 2
    xf, outf, _ = core.solve(h, b, max_iters=500)
    assert outf["status"] == 1
    x2, out2, _ = core.solve(h, b, max_iters=outf["iters"] + 1)
    assert out2["iters"] == outf["iters"] and out2["status"] == 1
    assert np.array_equal(x2, xf)
    # zero right-hand side: converges in the first body with x = 0
    xz, outz, _ = core.solve(h, np.zeros_like(b), max_iters=500)
    assert outz["iters"] == 0 and outz["status"] == 1 and not np.any(xz)
