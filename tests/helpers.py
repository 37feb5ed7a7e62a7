"""Shared test helpers: random hierarchies and an independent matrix assembly."""
import numpy as np
import scipy.sparse as sp

FLUID, AIR, SOLID = 0, 1, 2


def level_shapes(res, L):
    return [tuple(n >> l for n in res) for l in range(L)]


def random_hierarchy(res, L, seed, dtype=np.float32, p_fluid=0.7, p_air=0.2, sym_pos=True,
                     coupling=None):
    """Random cell types (coarsened by the oracle rule), random negative
    couplings between FLUID neighbours (or the single value `coupling`, as
    the pressure operators have) and a diagonally dominant Adiag."""
    from oracle import core

    rng = np.random.default_rng(seed)
    dim = len(res)
    u = rng.uniform(size=res)
    ct = np.where(u < p_fluid, FLUID, np.where(u < p_fluid + p_air, AIR, SOLID)).astype(np.int8)
    gts = [ct]
    for _ in range(1, L):
        gts.append(core.coarsen(gts[-1]))
    Ad, Ax = [], []
    for gt in gts:
        shape = gt.shape
        ax = np.zeros(shape + (dim,), dtype)
        fl = gt == FLUID
        for k in range(dim):
            nb = np.zeros(shape, bool)
            sl_lo = [slice(None)] * dim
            sl_hi = [slice(None)] * dim
            sl_lo[k] = slice(0, shape[k] - 1)
            sl_hi[k] = slice(1, shape[k])
            nb[tuple(sl_lo)] = fl[tuple(sl_lo)] & fl[tuple(sl_hi)]
            val = -rng.uniform(0.5, 1.5, shape) if coupling is None else coupling
            ax[..., k] = np.where(nb, val, 0.0).astype(dtype)
        off = np.zeros(shape, np.float64)
        for k in range(dim):
            off += np.abs(ax[..., k])
            sl_lo = [slice(None)] * dim
            sl_hi = [slice(None)] * dim
            sl_lo[k] = slice(0, shape[k] - 1)
            sl_hi[k] = slice(1, shape[k])
            tmp = np.zeros(shape)
            tmp[tuple(sl_hi)] = np.abs(ax[..., k])[tuple(sl_lo)]
            off += tmp
        ad = (off + rng.uniform(0.1, 1.0, shape)).astype(dtype)
        ad = np.where(fl, ad, 0.0).astype(dtype)
        Ad.append(np.ascontiguousarray(ad))
        Ax.append(np.ascontiguousarray(ax))
    return {"gt": gts, "Adiag": Ad, "Ax": Ax}


def assemble_matrix(gt, Adiag, Ax, fluid_only=True):
    """scipy.sparse A with A[I,I] = Adiag[I], A[I,I+e_k] = A[I+e_k,I] = Ax[I][k]
    (the operator neighbor_sum + Adiag applies, mgpcg.py:96-102, 218-224),
    restricted to FLUID rows/columns when fluid_only."""
    shape = gt.shape
    dim = len(shape)
    N = gt.size
    idx = np.arange(N).reshape(shape)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [Adiag.astype(np.float64).ravel()]
    for k in range(dim):
        sl_lo = [slice(None)] * dim
        sl_hi = [slice(None)] * dim
        sl_lo[k] = slice(0, shape[k] - 1)
        sl_hi[k] = slice(1, shape[k])
        a = idx[tuple(sl_lo)].ravel()
        b = idx[tuple(sl_hi)].ravel()
        v = Ax[..., k][tuple(sl_lo)].astype(np.float64).ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [v, v]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    if fluid_only:
        f = (gt.ravel() == FLUID).astype(np.float64)
        D = sp.diags(f)
        A = D @ A @ D
    return A.tocsr()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
