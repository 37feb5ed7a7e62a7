"""The packed level layout against a numpy restatement of its rules.

For every loaded level, the per-slot flag words (nonzero-coupling bits and
the diagonal code) and the per-block FLUID flags that the packing kernel
writes are recomputed here from the canonical arrays and compared exactly:
- bit a: the coupling to the upper neighbour on original axis a is nonzero
  (its FP32 bit pattern, so -0.0 counts), bit 3 + a: the lower one;
- bits 6..8: the smallest n < 7 for which -c added n times to 0 in FP32
  reproduces Adiag bit for bit, c being the slot's first nonzero coupling
  (upper then lower, axis by axis; n = 0 needs no coupling), else 7;
- bit 9: the slot is FLUID;
- one flag per 256 slots: 1 when a slot of either colour is FLUID;
- the work orders: per block of 256 slots the natural order with the wet
  bits; per half-sweep virtual block of 512 slots (level 0, on the batched
  kernels) a permutation whose wet bits match the flags, with the wet virtual
  blocks spread evenly over the positions.
Runs on the MI355X (`dev` = "cuda:0") and on the CPU emulation build.
"""
import numpy as np
import pytest
import torch

from helpers import random_hierarchy

FLUID = 0


def _fold(n, c):
    a = np.float32(0.0)
    m = np.float32(-np.float32(c))
    for _ in range(n):
        a = np.float32(a + m)
    return a


def _bits(x):
    return np.float32(x).view(np.uint32)


def _expected(gt, Ad, Ax):
    """Flag words [2][slots] and block flags of one level (padded layout)."""
    dim = gt.ndim
    if dim == 2:
        gt, Ad, Ax = gt[None], Ad[None], Ax[None]
    n0, n1, n2 = gt.shape
    P0 = n0 + 2 if dim == 3 else 1
    P1, W = n1 + 2, (n2 + 3) // 2
    size = P0 * P1 * W
    flags = np.zeros((2, size), np.uint16)
    fluid = np.zeros((2, size), bool)
    axb = Ax.view(np.uint32)
    for i in range(n0):
        for j in range(n1):
            for k in range(n2):
                ip = i + 1 if dim == 3 else 0
                t = (ip * P1 + j + 1) * W + ((k + 1) >> 1)
                c = (i + j + k) & 1
                f, cl, have = 0, 0.0, False
                v = (i, j, k)
                for a in range(dim):
                    if axb[i, j, k, a] != 0:
                        f |= 1 << a
                        if not have:
                            cl, have = Ax[i, j, k, a], True
                    va = a + 3 - dim
                    if v[va] > 0:
                        w = list(v)
                        w[va] -= 1
                        if axb[w[0], w[1], w[2], a] != 0:
                            f |= 1 << (3 + a)
                            if not have:
                                cl, have = Ax[w[0], w[1], w[2], a], True
                code = 7
                for n in range(7):
                    if n > 0 and not have:
                        break
                    if _bits(_fold(n, cl)) == _bits(Ad[i, j, k]):
                        code = n
                        break
                fluid[c, t] = gt[i, j, k] == FLUID
                flags[c, t] = f | (code << 6) | (int(fluid[c, t]) << 9)
    nb = (size + 255) // 256
    pad = np.zeros((2, nb * 256), bool)
    pad[:, :size] = fluid
    wet = pad.reshape(2, nb, 256).any(axis=(0, 2)).astype(np.uint8)
    return flags, wet, fluid


def _check_order(o, wetv, what):
    """o: a work order over len(wetv) entries whose FLUID flags are wetv."""
    n = len(wetv)
    blk = o & 0x7FFFFFFF
    isw = (o >> 31) & 1
    assert len(o) == n, what
    assert np.array_equal(np.sort(blk), np.arange(n)), f"{what}: not a permutation"
    assert np.array_equal(isw, wetv[blk]), f"{what}: wet bit"
    # wet entries, and dry ones, keep ascending block order
    assert np.all(np.diff(blk[isw == 1]) > 0) and np.all(np.diff(blk[isw == 0]) > 0), what
    # evenly spread: every prefix holds its proportional share of wet entries, +-1
    W = int(wetv.sum())
    pref = np.concatenate([[0], np.cumsum(isw)])
    share = np.arange(n + 1) * W / n
    assert np.all(np.abs(pref - share) < 1.0), f"{what}: wet positions not spread"


def _check(s, gts, Ads, Axs):
    for l in range(len(gts)):
        flags, wet = s.level_flags(l)
        ef, ew, _ = _expected(gts[l], Ads[l], Axs[l])
        assert np.array_equal(flags, ef), f"level {l} flag words"
        assert np.array_equal(wet, ew), f"level {l} block flags"
        o1, o2 = s.level_order(l)
        assert np.array_equal(o1 & 0x7FFFFFFF, np.arange(len(ew))), f"level {l} block order"
        assert np.array_equal(o1 >> 31, ew), f"level {l} block order wet bits"
        if o2 is None:
            assert l > 0 or s.info(0)["coarse_from"] == 0
            continue
        w2 = np.zeros(2 * len(o2), np.int64)
        w2[:len(ew)] = ew
        _check_order(o2, w2.reshape(-1, 2).max(axis=1), f"level {l} half-sweep order")


@pytest.mark.parametrize("coupling", [None, -0.75], ids=["random", "uniform"])
@pytest.mark.parametrize("res,L", [((32, 24), 3), ((16, 12, 10), 2)], ids=["2d", "3d"])
def test_flags_of_random_hierarchy(dev, res, L, coupling):
    from mgpcg_amd import mgpcg

    h = random_hierarchy(list(res), L, seed=21, coupling=coupling)
    s = mgpcg.MGPCGPoissonSolver(len(res), list(res), L, 2, 10, precision="plain", device=dev)
    for l in range(L):
        s.grid_type[l].copy_(torch.from_numpy(h["gt"][l]))
        s.Adiag[l].copy_(torch.from_numpy(h["Adiag"][l].astype(np.float32)))
        s.Ax[l].copy_(torch.from_numpy(h["Ax"][l].astype(np.float32)))
    s.load()
    _check(s, h["gt"], [a.astype(np.float32) for a in h["Adiag"]],
           [a.astype(np.float32) for a in h["Ax"]])


def test_flags_of_pressure_scene(dev):
    """A ghost-fluid pool+sphere pressure system: the level's operator as the
    strategy assembled it, and most FLUID cells get a decoded diagonal."""
    from mgpcg_amd import mgpcg
    from mgpcg_amd import scenes
    from mgpcg_amd import systems

    sysd = scenes.c2_3d(seed=1, n=16, L=3)
    s = mgpcg.MGPCGPoissonSolver(3, sysd["res"], 3, 2, 10, precision="mixed", device=dev)
    systems.assemble(s, systems.upload(sysd, dev))
    s.load()
    gts = [g.cpu().numpy() for g in s.grid_type]
    Ads = [a.cpu().numpy() for a in s.Adiag]
    Axs = [a.cpu().numpy() for a in s.Ax]
    _check(s, gts, Ads, Axs)
    flags, wet = s.level_flags(0)
    _, _, fluid = _expected(gts[0], Ads[0], Axs[0])
    codes = (flags >> 6) & 7
    assert fluid.any() and wet.any()
    # FLUID cells away from the free surface have plain-sum diagonals
    assert (codes[fluid] < 7).mean() > 0.5


@pytest.mark.parametrize("precision", ["plain", "mixed", "f64"])
def test_pool_bytes_is_what_create_allocates(dev, precision):
    """The admission charge comes from the library's own layout: the size
    query equals the bytes the created handle holds, for 2D and 3D shapes."""
    from mgpcg_amd import mgpcg

    for res, L in (([16, 16, 32], 3), ([32, 48, 16], 4), ([64, 32], 4), ([16, 16], 1)):
        q = mgpcg.pool_bytes(res, L, precision)
        s = mgpcg.MGPCGPoissonSolver(len(res), res, L, 2, 10, precision=precision, device=dev)
        assert s.info(0)["device_bytes"] == q, (res, L)
        canon = sum(t.numel() * t.element_size()
                    for t in [s.b, s.x] + s.grid_type + s.Adiag + s.Ax)
        if precision != "plain":
            canon += s.x.numel() * 8  # x64, allocated at the first readback
        assert mgpcg.resident_bytes(res, L, precision) == q + canon
        del s
