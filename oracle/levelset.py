"""numpy restatement of cell classification and fast-sweeping redistancing --
TEST INFRASTRUCTURE ONLY.

  apply_markers   fluid_simulator.py:151-159
  init_boundary   fluid_simulator.py:337-341
  redistance      FastSweepingLevelSet (level_set.py:260-333): target_surface
                  (level_set.py:66-86) and the Eikonal update
                  update_from_neighbor (level_set.py:88-112), FP32.
  extrap_velocity fluid_simulator.py:240-270 (faces outside a component's
                  own shape count as invalid).

Sweep semantics (shared with the HIP kernels; the reference's parallel lines
race): along the sweep axis a line is walked in order, so cells behind are
already updated and cells ahead are not; the neighbouring lines are read as
they were before the sweep.  Lines are therefore independent, and each
position along the axis is evaluated for all lines at once here.  Every axis
is swept over its own extent.
"""
import numpy as np

FLUID, AIR, SOLID = 0, 1, 2
f32 = np.float32


def apply_markers(cell_type, phi):
    out = cell_type.copy()
    free = out != SOLID
    out[free] = AIR
    out[free & (phi <= f32(0.0))] = FLUID
    return out


def init_boundary(cell_type):
    out = cell_type.copy()
    for a in range(out.ndim):
        lo = [slice(None)] * out.ndim
        hi = [slice(None)] * out.ndim
        lo[a] = 0
        hi[a] = -1
        out[tuple(lo)] = SOLID
        out[tuple(hi)] = SOLID
    return out


def _sign(x):
    return np.sign(x).astype(f32)  # -1, 0, 1 as taichi_glsl's sign


def target_surface(phi, dx):
    """-> (phi_temp, fixed)."""
    dim = phi.ndim
    dxf = f32(dx)
    est = np.full(phi.shape, f32(1e20), f32)
    change = np.zeros(phi.shape, bool)
    for k in range(dim):
        n = phi.shape[k]
        for s in (-1, 1):
            # neighbour I + s e_k where it exists
            src = [slice(None)] * dim
            dst = [slice(None)] * dim
            if s > 0:
                src[k], dst[k] = slice(1, n), slice(0, n - 1)
            else:
                src[k], dst[k] = slice(0, n - 1), slice(1, n)
            pi = phi[tuple(dst)]
            pj = phi[tuple(src)]
            sc = _sign(pi) != _sign(pj)
            with np.errstate(divide="ignore", invalid="ignore"):
                th = (pi / (pi - pj)).astype(f32)
            est0 = ((_sign(pi) * th).astype(f32) * dxf).astype(f32)
            e = est[tuple(dst)]
            take = sc & (np.abs(est0) < np.abs(e))
            est[tuple(dst)] = np.where(take, est0, e)
            change[tuple(dst)] |= sc
    tmp = np.where(change, est, f32(1e20)).astype(f32)
    return tmp, change


def _eikonal(nbs, dx, dim):
    """update_from_neighbor's closed form on sorted neighbour values."""
    dxf = f32(dx)
    two_dx2 = f32(2.0 * (dx * dx))
    dx2 = f32(dx * dx)
    half, third = f32(1.0 / 2.0), f32(1.0 / 3.0)
    nb = [x.copy() for x in nbs]
    for i in range(dim - 1):
        for j in range(dim - 1 - i):
            sw = nb[j] > nb[j + 1]
            a, b = nb[j].copy(), nb[j + 1].copy()
            nb[j] = np.where(sw, b, a)
            nb[j + 1] = np.where(sw, a, b)
    d = (nb[0] + dxf).astype(f32)
    c1 = d > nb[1]
    df = (nb[1] - nb[0]).astype(f32)
    # The two- and three-neighbour forms are evaluated on every entry and
    # selected where they apply; where they do not, their operands may be the
    # 1e20 "unknown" value, whose squares overflow.  Those entries are
    # discarded, and the selected result is checked to be finite.
    with np.errstate(over="ignore", invalid="ignore"):
        d2 = (half * ((nb[0] + nb[1]).astype(f32) +
                      np.sqrt((two_dx2 - (df * df).astype(f32)).astype(f32)))).astype(f32)
        d = np.where(c1, d2, d)
        if dim == 3:
            c2 = c1 & (d > nb[2])
            s = ((nb[0] + nb[1]).astype(f32) + nb[2]).astype(f32)
            q = (((nb[0] * nb[0]).astype(f32) + (nb[1] * nb[1]).astype(f32)).astype(f32) +
                 (nb[2] * nb[2]).astype(f32)).astype(f32)
            q = (q - dx2).astype(f32)
            inner = np.maximum(f32(0.0), ((s * s).astype(f32) - (f32(3.0) * q).astype(f32))
                               .astype(f32))
            d3 = (third * (s + np.sqrt(inner).astype(f32)).astype(f32)).astype(f32)
            d = np.where(c2, d3, d)
    assert np.all(np.isfinite(d)), "Eikonal update produced a non-finite distance"
    return d.astype(f32)


def _sweep(phi, tmp, fixed, dx, ax, direction):
    """One directional sweep along axis ``ax`` (updates ``tmp`` in place)."""
    dim = phi.ndim
    snap = tmp.copy()
    n = phi.shape[ax]
    T = np.moveaxis(tmp, ax, 0)  # views: writing T writes tmp
    Sn = np.moveaxis(snap, ax, 0)
    P = np.moveaxis(phi, ax, 0)
    F = np.moveaxis(fixed, ax, 0)
    other = [a for a in range(dim) if a != ax]
    order = range(n) if direction > 0 else range(n - 1, -1, -1)
    for p in order:
        nbs = []
        for k in range(dim):
            nk = phi.shape[k]
            if k == ax:
                up = T[p + 1] if p + 1 < n else T[p]  # the edge value is never selected
                dn = T[p - 1] if p - 1 >= 0 else T[p]
                coord = np.full(T[p].shape, p)
            else:
                kk = other.index(k)
                up = np.roll(Sn[p], -1, axis=kk)  # wrapped entries are never selected
                dn = np.roll(Sn[p], 1, axis=kk)
                shp = [1] * (dim - 1)
                shp[kk] = nk
                coord = np.broadcast_to(np.arange(nk).reshape(shp), T[p].shape)
            use_up = (coord == 0) | ((coord < nk - 1) & (np.abs(up) < np.abs(dn)))
            nbs.append(np.where(use_up, np.abs(up), np.abs(dn)).astype(f32))
        d = _eikonal(nbs, dx, dim)
        cur = T[p].copy()
        upd = (~F[p]) & (np.abs(d) < np.abs(cur))
        T[p] = np.where(upd, (d * _sign(P[p])).astype(f32), cur)
    return tmp


def redistance(phi, dx, repeat_times=2):
    """FastSweepingLevelSet.redistance (level_set.py:329-333) -> new phi."""
    phi = np.ascontiguousarray(phi, f32)
    tmp, fixed = target_surface(phi, dx)
    for _ in range(repeat_times):
        for ax in range(phi.ndim - 1, -1, -1):
            for direction in (1, -1):
                tmp = _sweep(phi, tmp, fixed, dx, ax, direction)
    return tmp


def extrap_velocity(cell_type, velocity, iterations=10):
    """-> new velocity list (mark_valid, then `iterations` diffuse passes)."""
    dim = cell_type.ndim
    out = []
    for k in range(dim):
        u = np.array(velocity[k], dtype=f32, copy=True)
        shape = u.shape
        # mark_valid: a cell on either side (I - e_k, I) is FLUID
        fl = np.zeros([n + 2 for n in cell_type.shape], bool)  # padded: out of range = no
        fl[tuple(slice(1, -1) for _ in range(dim))] = cell_type == FLUID
        # padded coordinates are I + 1: cell I - e_k, then cell I
        lo = [slice(1, n + 1) if a != k else slice(0, n) for a, n in enumerate(shape)]
        hi = [slice(1, n + 1) for n in shape]
        valid = (fl[tuple(lo)] | fl[tuple(hi)]).astype(np.int8)
        for _ in range(iterations):
            src = u.copy()
            vt = valid.copy()
            vp = np.zeros([n + 2 for n in shape], np.int8)
            sp = np.zeros([n + 2 for n in shape], f32)
            vp[tuple(slice(1, -1) for _ in range(dim))] = vt
            sp[tuple(slice(1, -1) for _ in range(dim))] = src
            tot = np.zeros(shape, f32)
            cnt = np.zeros(shape, np.int32)
            for off in np.ndindex(*([3] * dim)):  # C order over offsets -1..1
                sl = tuple(slice(o, o + n) for o, n in zip(off, shape))
                m = vp[sl] == 1
                tot = np.where(m, (tot + sp[sl]).astype(f32), tot)
                cnt = cnt + m
            upd = (vt == 0) & (cnt > 0)
            with np.errstate(invalid="ignore", divide="ignore"):
                avg = (tot / cnt.astype(f32)).astype(f32)
            u = np.where(upd, avg, u).astype(f32)
            valid = np.where(upd, np.int8(1), valid).astype(np.int8)
        out.append(u)
    return out

