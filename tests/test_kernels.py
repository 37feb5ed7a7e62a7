"""Kernel-level parity of the HIP library against the oracle.

Each stage of the V-cycle and the CG operator runs once through the C-ABI
(mgpcg_debug_set / _op / _get) on seeded inputs and is compared with the
oracle's restatement of the same reference kernel.  FP32 results must be bit
identical: both sides use the same operation order without contraction.
Every test runs on the MI355X (`dev` = "cuda:0", marked gpu) and on the CPU
emulation build of the same sources (`dev` = "cpu").
"""
import numpy as np
import pytest
import torch

from helpers import random_hierarchy
from oracle import core

CASES = [
    ((32, 24), 3),       # 2D
    ((16, 16, 16), 3),   # 3D
    ((48, 32, 16), 5),   # 3D, odd and length-1 coarse axes (3 x 2 x 1)
    ((24, 40, 8), 4),    # 3D, odd coarse axes (3 x 5 x 1)
]


def _solver(h, dev, precision="plain"):
    from mgpcg_amd import mgpcg

    res = list(h["gt"][0].shape)
    dim = len(res)
    L = len(h["gt"])
    s = mgpcg.MGPCGPoissonSolver(dim, res, L, 2, 10, precision=precision, device=dev)
    for l in range(L):
        s.grid_type[l].copy_(torch.from_numpy(h["gt"][l]))
        s.Adiag[l].copy_(torch.from_numpy(h["Adiag"][l].astype(np.float32)))
        s.Ax[l].copy_(torch.from_numpy(h["Ax"][l].astype(np.float32)))
    s.load()
    return s


def _get(s, field, level):
    return s.debug_get(field, level).cpu().numpy()


def _case_id(c):
    return "x".join(map(str, c[0])) + f"_L{c[1]}"


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_coarsen_bit_exact(dev, case):
    from mgpcg_amd import _lib

    res, _ = case
    rng = np.random.default_rng(1)
    fine = rng.integers(0, 3, res).astype(np.int8)
    t = torch.from_numpy(fine).to(dev)
    out = torch.empty([n // 2 for n in res], dtype=torch.int8, device=dev)
    lib = _lib.load()
    _lib.check(lib.mgpcg_coarsen(_lib.stream_handle(dev), len(res), _lib.int3(res), _lib.ptr(t),
                                 _lib.ptr(out)))
    _lib.synchronize(dev)
    assert np.array_equal(out.cpu().numpy(), core.coarsen(fine))


@pytest.mark.parametrize("coupling", [None, -0.75], ids=["random", "uniform"])
@pytest.mark.parametrize("case", CASES, ids=_case_id)
@pytest.mark.parametrize("phase", [0, 1])
def test_smooth_bit_exact(dev, case, phase, coupling):
    from mgpcg_amd import _lib

    res, L = case
    h = random_hierarchy(res, L, seed=2, coupling=coupling)
    s = _solver(h, dev)
    rng = np.random.default_rng(3)
    for l in range(L):
        shape = h["gt"][l].shape
        r = rng.standard_normal(shape).astype(np.float32)
        z = rng.standard_normal(shape).astype(np.float32)
        s.debug_set(_lib.FIELD_R, l, torch.from_numpy(r))
        s.debug_set(_lib.FIELD_Z, l, torch.from_numpy(z))
        s.debug_op(_lib.OP_SMOOTH, l, phase)
        got = _get(s, _lib.FIELD_Z, l)
        core.smooth(h["gt"][l], h["Adiag"][l], h["Ax"][l], r, z, phase)
        assert np.array_equal(got, z), f"level {l}"


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_smooth_pair_from_zero(dev, case):
    """z.fill(0); smooth(l, 0); smooth(l, 1) in one kernel (colour 0 recomputed
    on the fly inside the colour-1 stencil)."""
    from mgpcg_amd import _lib

    res, L = case
    h = random_hierarchy(res, L, seed=4)
    s = _solver(h, dev)
    rng = np.random.default_rng(5)
    for l in range(L):
        shape = h["gt"][l].shape
        r = rng.standard_normal(shape).astype(np.float32)
        s.debug_set(_lib.FIELD_R, l, torch.from_numpy(r))
        s.debug_set(_lib.FIELD_Z, l, torch.from_numpy(rng.standard_normal(shape).astype(np.float32)))
        s.debug_op(_lib.OP_SMOOTH_PAIR_FROM_ZERO, l, 0)
        got = _get(s, _lib.FIELD_Z, l)
        z = np.zeros(shape, np.float32)
        core.smooth(h["gt"][l], h["Adiag"][l], h["Ax"][l], r, z, 0)
        core.smooth(h["gt"][l], h["Adiag"][l], h["Ax"][l], r, z, 1)
        assert np.array_equal(got, z), f"level {l}"  # == treats -0 and +0 alike


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_prolong_smooth1_bit_exact(dev, case):
    """prolongate(l); smooth(l, 1) fused: every colour-1 FLUID value matches."""
    from mgpcg_amd import _lib

    res, L = case
    h = random_hierarchy(res, L, seed=16)
    s = _solver(h, dev)
    rng = np.random.default_rng(17)
    for l in range(L - 1):
        shape = h["gt"][l].shape
        r = rng.standard_normal(shape).astype(np.float32)
        z = rng.standard_normal(shape).astype(np.float32)
        zc = rng.standard_normal(h["gt"][l + 1].shape).astype(np.float32)
        s.debug_set(_lib.FIELD_R, l, torch.from_numpy(r))
        s.debug_set(_lib.FIELD_Z, l, torch.from_numpy(z))
        s.debug_set(_lib.FIELD_Z, l + 1, torch.from_numpy(zc))
        s.debug_op(_lib.OP_PROLONG_SMOOTH1, l, 0)
        got = _get(s, _lib.FIELD_Z, l)
        zr = z.copy()
        core.prolongate(zr, zc)
        core.smooth(h["gt"][l], h["Adiag"][l], h["Ax"][l], r, zr, 1)
        colour = np.indices(shape).sum(axis=0) & 1
        upd = (h["gt"][l] == 0) & (colour == 1)
        assert np.array_equal(got[upd], zr[upd]), f"level {l}"


@pytest.mark.parametrize("rows", [2, 1, 0], ids=["gather", "rows", "children"])
@pytest.mark.parametrize("coupling", [None, -0.75], ids=["random", "uniform"])
@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_restrict_bit_exact(dev, case, coupling, rows):
    """Every restriction kernel (residual + gather, row-tiled, child-parallel), with stored and
    with uniform (flag-word) couplings, against the oracle bit for bit.  The
    gather skips blocks of coarse cells none of which is FLUID (their r is
    never read), so it is compared on the FLUID coarse cells; the other two
    write every coarse cell and are compared everywhere."""
    from mgpcg_amd import _lib

    res, L = case
    h = random_hierarchy(res, L, seed=6, coupling=coupling)
    s = _solver(h, dev)
    lib = _lib.load()
    _lib.check(lib.mgpcg_set_option(None, _lib.OPT_RESTRICT_ROWS, rows))
    try:
        rng = np.random.default_rng(7)
        compared = 0
        for l in range(L - 1):
            if coupling is not None:
                assert s.level_coupling(l) == (True, np.float32(coupling)) or \
                    not np.any(h["Ax"][l])
            shape = h["gt"][l].shape
            r = rng.standard_normal(shape).astype(np.float32)
            z = rng.standard_normal(shape).astype(np.float32)
            s.debug_set(_lib.FIELD_R, l, torch.from_numpy(r))
            s.debug_set(_lib.FIELD_Z, l, torch.from_numpy(z))
            s.debug_op(_lib.OP_RESTRICT, l, 0)
            got = _get(s, _lib.FIELD_R, l + 1)
            ref = core.restrict(h["gt"][l], h["Adiag"][l], h["Ax"][l], r, z)
            if rows == 2:
                fl = h["gt"][l + 1] == 0
                compared += int(fl.sum())
                assert np.array_equal(got[fl], ref[fl]), f"level {l}"
            else:
                compared += got.size
                assert np.array_equal(got, ref), f"level {l}"
        assert compared > 0
    finally:
        _lib.check(lib.mgpcg_set_option(None, _lib.OPT_RESTRICT_ROWS, 2))


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_prolongate_bit_exact(dev, case):
    from mgpcg_amd import _lib

    res, L = case
    h = random_hierarchy(res, L, seed=8)
    s = _solver(h, dev)
    rng = np.random.default_rng(9)
    for l in range(L - 1):
        shape = h["gt"][l].shape
        cshape = h["gt"][l + 1].shape
        z = rng.standard_normal(shape).astype(np.float32)
        zc = rng.standard_normal(cshape).astype(np.float32)
        s.debug_set(_lib.FIELD_Z, l, torch.from_numpy(z))
        s.debug_set(_lib.FIELD_Z, l + 1, torch.from_numpy(zc))
        s.debug_op(_lib.OP_PROLONGATE, l, 0)
        got = _get(s, _lib.FIELD_Z, l)
        core.prolongate(z, zc)
        assert np.array_equal(got, z), f"level {l}"


@pytest.mark.parametrize("case", CASES, ids=_case_id)
@pytest.mark.parametrize("coupling", [None, -0.75], ids=["random", "uniform"])
def test_update_p_compute_Ap_bit_exact(dev, case, coupling):
    """p_next = z + 0 * p (update_p with beta = 0), Ap = A p_next."""
    from mgpcg_amd import _lib

    res, L = case
    h = random_hierarchy(res, L, seed=10, coupling=coupling)
    s = _solver(h, dev)
    rng = np.random.default_rng(11)
    z = rng.standard_normal(res).astype(np.float32)
    p = rng.standard_normal(res).astype(np.float32)
    s.debug_set(_lib.FIELD_Z, 0, torch.from_numpy(z))
    s.debug_set(_lib.FIELD_P, 0, torch.from_numpy(p))
    s.debug_set(_lib.FIELD_AP, 0, torch.zeros(res))
    s.debug_op(_lib.OP_COMPUTE_AP, 0, 0)
    fl = h["gt"][0] == 0
    pn = _get(s, _lib.FIELD_P_NEXT, 0)
    assert np.array_equal(pn[fl], z[fl])
    pref = np.where(fl, z + np.float32(0.0) * p, p).astype(np.float32)
    ref = core.compute_Ap(h["gt"][0], h["Adiag"][0], h["Ax"][0], pref)
    got = _get(s, _lib.FIELD_AP, 0)
    assert np.array_equal(got[fl], ref[fl])


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_vcycle_bit_exact(dev, case):
    """Every FLUID value of z = M r; non-FLUID z is inert (unprolonged)."""
    from mgpcg_amd import _lib

    res, L = case
    h = random_hierarchy(res, L, seed=12)
    s = _solver(h, dev)
    r0 = np.random.default_rng(13).standard_normal(res).astype(np.float32)
    s.debug_set(_lib.FIELD_R, 0, torch.from_numpy(r0))
    s.debug_op(_lib.OP_VCYCLE, 0, 0)
    got = _get(s, _lib.FIELD_Z, 0)
    ref, _, _ = core.v_cycle(h, r0, nu=2, bottom=10)
    fl = h["gt"][0] == 0
    assert np.array_equal(got[fl], ref[fl])


@pytest.mark.parametrize("case", CASES[1:3], ids=_case_id)
@pytest.mark.parametrize("coarse_max", [0, 1 << 30])
def test_vcycle_schedule_variants(dev, case, coarse_max):
    """All levels on batched kernels (coarse_max 0) or all in the
    one-workgroup coarse kernel: the same FLUID values."""
    from mgpcg_amd import _lib

    res, L = case
    lib = _lib.load()
    _lib.check(lib.mgpcg_set_option(None, _lib.OPT_COARSE_MAX_SLOTS, coarse_max))
    try:
        h = random_hierarchy(res, L, seed=18)
        s = _solver(h, dev)
    finally:
        _lib.check(lib.mgpcg_set_option(None, _lib.OPT_COARSE_MAX_SLOTS, 4096))
    r0 = np.random.default_rng(19).standard_normal(res).astype(np.float32)
    s.debug_set(_lib.FIELD_R, 0, torch.from_numpy(r0))
    s.debug_op(_lib.OP_VCYCLE, 0, 0)
    got = _get(s, _lib.FIELD_Z, 0)
    ref, _, _ = core.v_cycle(h, r0, nu=2, bottom=10)
    fl = h["gt"][0] == 0
    assert np.array_equal(got[fl], ref[fl])


@pytest.mark.parametrize("case", CASES[:2], ids=_case_id)
def test_vcycle_f64_matches_oracle(dev, case):
    from mgpcg_amd import _lib

    res, L = case
    h = core.to_f64(random_hierarchy(res, L, seed=14))
    s = _solver(h, dev, precision="f64")
    r0 = np.random.default_rng(15).standard_normal(res)
    s.debug_set(_lib.FIELD_R, 0, torch.from_numpy(r0))
    s.debug_op(_lib.OP_VCYCLE, 0, 0)
    got = _get(s, _lib.FIELD_Z, 0)
    ref, _, _ = core.v_cycle(h, r0, nu=2, bottom=10)
    fl = h["gt"][0] == 0
    assert np.array_equal(got[fl], ref[fl])
