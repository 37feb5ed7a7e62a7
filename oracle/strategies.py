"""numpy restatements of the operator-assembly strategies -- TEST INFRASTRUCTURE ONLY.

Each function follows one reference kernel (file:line cited) in FP32 with the
reference's per-cell operation order.  Python-scope constants are converted
to FP32 before use, as Taichi bakes them into FP32 kernels (fluid_simulator.py:16
leaves ``default_fp`` at f32).

Conventions shared with the HIP strategies (the reference leaves these
undefined):
  * an out-of-range neighbour cell is SOLID (the reference reads
    ``grid_type[I + offset]`` out of bounds at coarse-level edges,
    pressure_project.py:55);
  * the ghost-fluid weight ``c = (phi_i - phi_j) / phi_i`` is replaced by the
    clamp value when ``phi_i == 0`` (the reference would store -inf,
    pressure_project.py:36-37,61-62); otherwise ``min(c, c_max)`` exactly as
    in the reference, with ``c_max = 1e3`` by default.
"""
import math

import numpy as np

FLUID, AIR, SOLID = 0, 1, 2
f32 = np.float32


# ----------------------------------------------------------------------------
# utils.sample (utils.py:12-32) and LevelSet.theta / delta (level_set.py:24-37)
# ----------------------------------------------------------------------------
def sample(data, pos):
    """Clamped multilinear interpolation of ``data`` at index-space ``pos``.

    ``pos`` is a tuple of FP32 arrays (one per axis).  i = clamp(int(pos)),
    int() truncating toward zero; ip = clamp(i + 1); s = clamp(pos - i, 0, 1).
    The blend order follows utils.py:19-21 (2D) and utils.py:27-32 (3D).
    """
    dim = data.ndim
    idx, idxp, w = [], [], []
    for a in range(dim):
        p = pos[a].astype(f32)
        n = data.shape[a]
        i = np.clip(np.trunc(p).astype(np.int64), 0, n - 1)
        ip = np.clip(i + 1, 0, n - 1)
        s = np.clip(p - i.astype(f32), f32(0.0), f32(1.0)).astype(f32)
        idx.append(i)
        idxp.append(ip)
        w.append(s)
    one = f32(1.0)
    if dim == 2:
        i, j = idx
        ip, jp = idxp
        s, t = w
        d = data
        return ((d[i, j] * (one - s) + d[ip, j] * s) * (one - t)
                + (d[i, jp] * (one - s) + d[ip, jp] * s) * t).astype(f32)
    i, j, k = idx
    ip, jp, kp = idxp
    s, t, u = w
    d = data
    return ((((d[i, j, k] * (one - s) + d[ip, j, k] * s) * (one - t)
              + (d[i, jp, k] * (one - s) + d[ip, jp, k] * s) * t) * (one - u))
            + (((d[i, j, kp] * (one - s) + d[ip, j, kp] * s) * (one - t)
                + (d[i, jp, kp] * (one - s) + d[ip, jp, kp] * s) * t) * u)).astype(f32)


def theta(phi, dx):
    """Smoothed Heaviside (level_set.py:24-30), eps = 5 dx (level_set.py:22)."""
    eps = 5 * dx
    phi = phi.astype(f32)
    mid = (f32(0.5) + phi / f32(2 * eps)) + f32(1 / (2 * math.pi)) * np.sin(
        f32(math.pi) * phi / f32(eps)).astype(f32)
    out = np.where(phi <= f32(-eps), f32(0.0), np.where(phi >= f32(eps), f32(1.0), mid))
    return out.astype(f32)


def delta(phi, dx):
    """Smoothed Dirac delta (level_set.py:33-37), eps = 5 dx."""
    eps = 5 * dx
    phi = phi.astype(f32)
    mid = (f32(1.0) + np.cos(f32(math.pi) * phi / f32(eps)).astype(f32)) / f32(2 * eps)
    out = np.where((phi <= f32(-eps)) | (phi >= f32(eps)), f32(0.0), mid)
    return out.astype(f32)


def _grid(shape):
    return np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")


def _shift(a, axis, s, fill):
    """out[I] = a[I + s e_axis], ``fill`` where out of range."""
    out = np.full_like(a, fill)
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    if s > 0:
        src[axis] = slice(s, None)
        dst[axis] = slice(None, -s)
    else:
        src[axis] = slice(None, s)
        dst[axis] = slice(-s, None)
    out[tuple(dst)] = a[tuple(src)]
    return out


def _gfm_c(phi_i, phi_j, c_max):
    with np.errstate(divide="ignore", invalid="ignore"):
        c = ((phi_i - phi_j) / phi_i).astype(f32)
    c = np.minimum(c, f32(c_max))
    return np.where(phi_i == f32(0.0), f32(c_max), c).astype(f32)


# ----------------------------------------------------------------------------
# PressureProjectStrategy (pressure_project.py:6-70)
# ----------------------------------------------------------------------------
def pressure_build_b(cell_type, velocity, scale_A, scale_b, p0, ghost_fluid_method=True,
                     phi=None, c_max=1e3, air_term=None):
    """build_b_kernel (pressure_project.py:15-39).

    ``air_term`` (optional callable (k, s) -> FP32 array) replaces the AIR-face
    term; MagneticForceStrategy uses it (magnetic_force.py:34-38).
    """
    dim = cell_type.ndim
    sA, sb, p0f = f32(scale_A), f32(scale_b), f32(p0)
    fluid = cell_type == FLUID
    b = np.zeros(cell_type.shape, f32)
    acc = np.zeros(cell_type.shape, f32)
    sl = tuple(slice(0, n) for n in cell_type.shape)
    for k in range(dim):
        u = velocity[k]
        lo = u[sl]
        hi_sl = list(sl)
        hi_sl[k] = slice(1, cell_type.shape[k] + 1)
        hi = u[tuple(hi_sl)]
        acc = (acc + (lo - hi)).astype(f32)
    b = np.where(fluid, (acc * sb).astype(f32), b)
    for k in range(dim):
        u = velocity[k]
        lo = u[sl]
        hi_sl = list(sl)
        hi_sl[k] = slice(1, cell_type.shape[k] + 1)
        hi = u[tuple(hi_sl)]
        for s in (-1, 1):
            nb = _shift(cell_type, k, s, SOLID)
            solid = fluid & (nb == SOLID)
            air = fluid & (nb == AIR)
            if s < 0:
                b = np.where(solid, (b - sb * (lo - f32(0.0))).astype(f32), b)
            else:
                b = np.where(solid, (b + sb * (hi - f32(0.0))).astype(f32), b)
            if air_term is not None:
                term = air_term(k, s)
            elif ghost_fluid_method:
                c = _gfm_c(phi, _shift(phi, k, s, f32(1.0)), c_max)
                term = ((sA * c).astype(f32) * p0f).astype(f32)
            else:
                term = np.full(cell_type.shape, (sA * p0f), f32)
            b = np.where(air, (b + term).astype(f32), b)
    return b.astype(f32)


def pressure_build_A(grid_type, level, scale_A, ghost_fluid_method=True, phi=None, c_max=1e3):
    """build_A_kernel (pressure_project.py:45-64) at one level -> (Adiag, Ax)."""
    dim = grid_type.ndim
    sA = f32(scale_A)
    fluid = grid_type == FLUID
    Adiag = np.zeros(grid_type.shape, f32)
    Ax = np.zeros(grid_type.shape + (dim,), f32)
    for k in range(dim):
        for s in (-1, 1):
            nb = _shift(grid_type, k, s, SOLID)
            m_f = fluid & (nb == FLUID)
            m_a = fluid & (nb == AIR)
            Adiag = np.where(m_f, (Adiag + sA).astype(f32), Adiag)
            if s > 0:
                Ax[..., k] = np.where(m_f, -sA, Ax[..., k])
            if ghost_fluid_method and level == 0:
                c = _gfm_c(phi, _shift(phi, k, s, f32(1.0)), c_max)
                Adiag = np.where(m_a, (Adiag + (sA * c).astype(f32)).astype(f32), Adiag)
            else:
                Adiag = np.where(m_a, (Adiag + sA).astype(f32), Adiag)
    return Adiag, Ax


# ----------------------------------------------------------------------------
# PressureProjectWithVolumeControlStrategy (volume_control.py:7-51)
# ----------------------------------------------------------------------------
def calc_volume(phi, dx):
    """calc_volume (volume_control.py:22-31), accumulated in FP64."""
    dim = phi.ndim
    I = _grid(phi.shape)
    tot = np.zeros(phi.shape, f32)
    for off in np.ndindex(*([2] * dim)):
        pos = tuple((I[a] + off[a]).astype(f32) - f32(0.5) for a in range(dim))
        tot = (tot + theta(-sample(phi, pos), dx)).astype(f32)
    cell = (tot / f32(2 ** dim)).astype(f32)
    return float(np.sum(cell.astype(np.float64)))


class VolumeController:
    """PI controller state of volume_control.py:13-21,39-51 (Python doubles)."""

    def __init__(self, dt, n_p=25, zeta=2):
        self.dt = dt
        self.step = 0
        self.vol_0 = 0.0
        self.y = 0.0
        self.k_p = 2.3 / (n_p * dt)
        self.k_i = (self.k_p / (2 * zeta)) ** 2

    def update(self, vol):
        """Returns the divergence constant c, or None on the first call."""
        self.step += 1
        if self.step == 1:
            self.vol_0 = vol
            return None
        x = (vol - self.vol_0) / self.vol_0
        self.y += x * self.dt
        return (-self.k_p * x - self.k_i * self.y) / (x + 1)


def add_c(cell_type, b, c):
    """add_c (volume_control.py:33-37): b[FLUID] += f32(c)."""
    return np.where(cell_type == FLUID, (b + f32(c)).astype(f32), b).astype(f32)


# ----------------------------------------------------------------------------
# MagneticForceStrategy.build_b_kernel (magnetic_force.py:17-39)
# ----------------------------------------------------------------------------
def H_interp(H, pos):
    """FerrofluidSimulator.H_interp (ferrofluid_simulator.py:66-70)."""
    dim = len(H)
    out = []
    for k in range(dim):
        p = tuple((pos[a] - f32(0.0 if a == k else 0.5)).astype(f32) for a in range(dim))
        out.append(sample(H[k], p))
    return out


def magnetic_build_b(cell_type, velocity, H, scale_A, scale_b, p0, mu0, kk):
    dim = cell_type.ndim
    I = _grid(cell_type.shape)
    sA, p0f = f32(scale_A), f32(p0)
    coef = f32(1 / 2 * kk * mu0)

    def air_term(k, s):
        pos = tuple((I[a] + (s if a == k else 0)).astype(f32) + f32(0.5) for a in range(dim))
        Hv = H_interp(H, pos)
        hh = (Hv[0] * Hv[0]).astype(f32)
        for a in range(1, dim):
            hh = (hh + Hv[a] * Hv[a]).astype(f32)
        return (sA * (p0f + (coef * hh).astype(f32))).astype(f32)

    return pressure_build_b(cell_type, velocity, scale_A, scale_b, p0, air_term=air_term)


# ----------------------------------------------------------------------------
# The step after the solve: apply_pressure (fluid_simulator.py:192-214) and
# apply_magnetic_force (ferrofluid_simulator.py:114-137)
# ----------------------------------------------------------------------------
def apply_pressure(cell_type, velocity, pressure, dt, rho, dx, p0, ghost_fluid_method=True,
                   phi=None, c_max=1e3, H=None, mu0=None, kk=None):
    """Returns the updated MAC velocity (a new list).  For each axis k and
    cell I, the face u_k[I] between I_1 = I - e_k and I changes when either
    side is FLUID: 0 next to SOLID (out of range is SOLID); fluid-air and
    air-fluid faces take p0 on the AIR side (ghost-fluid weighted), or with
    ``H`` given (apply_magnetic_force) p0 + 1/2 k mu0 |H|^2 at the AIR
    cell's centre and no weight; fluid-fluid faces the plain jump."""
    dim = cell_type.ndim
    scale = f32(f32(dt) / f32(rho * dx))
    p0f = f32(p0)
    I = _grid(cell_type.shape)
    out = [np.array(u, dtype=f32, copy=True) for u in velocity]
    sl = tuple(slice(0, n) for n in cell_type.shape)
    magnetic = H is not None
    if magnetic:
        coef = f32(1 / 2 * kk * mu0)

        def p_air(shift_k):
            pos = tuple((I[a] + (shift_k if a == k else 0)).astype(f32) + f32(0.5)
                        for a in range(dim))
            Hv = H_interp(H, pos)
            hh = (Hv[0] * Hv[0]).astype(f32)
            for a in range(1, dim):
                hh = (hh + Hv[a] * Hv[a]).astype(f32)
            return (p0f + (coef * hh).astype(f32)).astype(f32)

    for k in range(dim):
        u = out[k][sl].copy()
        ti, tj = cell_type, _shift(cell_type, k, -1, SOLID)
        pi, pj = pressure.astype(f32), _shift(pressure.astype(f32), k, -1, f32(0.0))
        act = (tj == FLUID) | (ti == FLUID)
        solid = act & ((tj == SOLID) | (ti == SOLID))
        fa = act & ~solid & (ti == AIR)
        af = act & ~solid & ~fa & (tj == AIR)
        ff = act & ~solid & ~fa & ~af
        new = u.copy()
        new[solid] = f32(0.0)
        if magnetic:
            v_fa = (u - (scale * (p_air(0) - pj)).astype(f32)).astype(f32)
            v_af = (u - (scale * (pi - p_air(-1))).astype(f32)).astype(f32)
        elif ghost_fluid_method:
            phj = _shift(phi, k, -1, f32(1.0))
            c_fa = _gfm_c(phj, phi, c_max)
            c_af = _gfm_c(phi, phj, c_max)
            v_fa = (u - ((scale * (p0f - pj)).astype(f32) * c_fa).astype(f32)).astype(f32)
            v_af = (u - ((scale * (pi - p0f)).astype(f32) * c_af).astype(f32)).astype(f32)
        else:
            v_fa = (u - (scale * (p0f - pj)).astype(f32)).astype(f32)
            v_af = (u - (scale * (pi - p0f)).astype(f32)).astype(f32)
        v_ff = (u - (scale * (pi - pj)).astype(f32)).astype(f32)
        new = np.where(fa, v_fa, new)
        new = np.where(af, v_af, new)
        new = np.where(ff, v_ff, new)
        out[k][sl] = new
    return out


def apply_magnetic_force(cell_type, velocity, pressure, H, dt, rho, dx, p0, mu0, kk):
    """apply_magnetic_force (ferrofluid_simulator.py:114-137)."""
    return apply_pressure(cell_type, velocity, pressure, dt, rho, dx, p0,
                          ghost_fluid_method=False, H=H, mu0=mu0, kk=kk)


# ----------------------------------------------------------------------------
# PotentialFunctionStrategy (potential_function.py:6-60)
# ----------------------------------------------------------------------------
def _in_domain(coords, res):
    m = np.ones(coords[0].shape, bool)
    for a, c in enumerate(coords):
        m &= (c >= 0) & (c + 1 <= res[a])
    return m


def potential_build_b(res, chi, H_ext):
    """build_b_kernel (potential_function.py:19-34)."""
    dim = len(res)
    I = _grid(res)
    b = np.zeros(res, f32)
    origin = np.ones(res, bool)
    for a in range(dim):
        origin &= I[a] == 0
    dom = _in_domain(I, res) & ~origin
    for k in range(dim):
        off = [f32(0.0 if a == k else 0.5) for a in range(dim)]
        up = [I[a] + (1 if a == k else 0) for a in range(dim)]
        dn = [I[a] - (1 if a == k else 0) for a in range(dim)]
        m_up = dom & _in_domain(up, res)
        pos_up = tuple(up[a].astype(f32) + off[a] for a in range(dim))
        Hk = H_ext[k]
        up_idx = tuple(np.clip(up[a], 0, Hk.shape[a] - 1) for a in range(dim))
        term = (sample(chi, pos_up) * Hk[up_idx]).astype(f32)
        b = np.where(m_up, (b + term).astype(f32), b)
        m_dn = dom & _in_domain(dn, res)
        pos_c = tuple(I[a].astype(f32) + off[a] for a in range(dim))
        term = (sample(chi, pos_c) * Hk[tuple(I)]).astype(f32)
        b = np.where(m_dn, (b - term).astype(f32), b)
    return np.where(origin, f32(0.0), b).astype(f32)


def potential_build_A(res, level, dx, chi):
    """build_A_kernel (potential_function.py:40-55) at one level -> (Adiag, Ax).

    ``in_domain`` tests the fine ``res`` at every level (potential_function.py:17).
    """
    dim = len(res)
    shape = tuple(n >> level for n in res)
    I = _grid(shape)
    # ``dx`` is a kernel-scope f32 variable in the reference (potential_function.py:43),
    # so 1 / dx and (1 + chi) / dx are f32 divisions by f32(dx_l)
    dxl = dx * (2 ** level)
    dxf = f32(dxl)
    inv = f32(1.0) / dxf
    Adiag = np.zeros(shape, f32)
    Ax = np.zeros(shape + (dim,), f32)
    origin = np.ones(shape, bool)
    for a in range(dim):
        origin &= I[a] == 0
    dom = _in_domain(I, res) & ~origin
    scale = f32(2 ** level)
    for k in range(dim):
        for s in (-1, 1):
            nbc = [I[a] + (s if a == k else 0) for a in range(dim)]
            m = dom & _in_domain(nbc, res)
            pos = tuple(((I[a].astype(f32) + f32(0.5 * s if a == k else 0.0)) * scale).astype(f32)
                        for a in range(dim))
            val = ((f32(1.0) + sample(chi, pos)) / dxf).astype(f32)
            Adiag = np.where(m, (Adiag - val).astype(f32), Adiag)
            if s > 0:
                Ax[..., k] = np.where(m, val, Ax[..., k])
    Adiag = np.where(origin, inv, Adiag).astype(f32)
    return Adiag, Ax


# ----------------------------------------------------------------------------
# SurfaceTension (surface_tension.py:95-153) and its strategy (:11-92)
# ----------------------------------------------------------------------------
def calc_n(phi, d, dx):
    """calc_n (surface_tension.py:115-127) -> n as [*res, dim]."""
    dim = phi.ndim
    I = _grid(phi.shape)
    off = [f32(0.0 if a == d else 0.5) for a in range(dim)]
    n = np.zeros(phi.shape + (dim,), f32)
    for k in range(dim):
        pp = tuple((I[a].astype(f32) + f32(0.5 if a == k else 0.0)) + off[a] for a in range(dim))
        pm = tuple((I[a].astype(f32) - f32(0.5 if a == k else 0.0)) + off[a] for a in range(dim))
        n[..., k] = ((sample(phi, pp) - sample(phi, pm)) / f32(dx)).astype(f32)
    nn = (n[..., 0] * n[..., 0]).astype(f32)
    for k in range(1, dim):
        nn = (nn + n[..., k] * n[..., k]).astype(f32)
    norm = np.sqrt(nn).astype(f32)
    safe = np.where(norm > 0, norm, f32(1.0))
    n = np.where((norm > 0)[..., None], (n / safe[..., None]).astype(f32), n)
    return n.astype(f32)


def calc_kappa(n, dx):
    """calc_kappa (surface_tension.py:129-136)."""
    dim = n.shape[-1]
    shape = n.shape[:-1]
    I = _grid(shape)
    kappa = np.zeros(shape, f32)
    for k in range(dim):
        pp = tuple(I[a].astype(f32) + f32(0.5 if a == k else 0.0) for a in range(dim))
        pm = tuple(I[a].astype(f32) - f32(0.5 if a == k else 0.0) for a in range(dim))
        diff = (sample(n[..., k], pp) - sample(n[..., k], pm)).astype(f32)
        kappa = (kappa + (diff / f32(dx)).astype(f32)).astype(f32)
    return kappa


def surface_build_b(velocity_d, phi, kappa, n, d, dt, dx, sigma=7.28e-2):
    """build_b_kernel (surface_tension.py:23-66) for velocity axis d."""
    dim = phi.ndim
    shape = phi.shape
    I = _grid(shape)
    u = velocity_d
    off = [f32(0.0 if a == d else 0.5) for a in range(dim)]
    idx = tuple(I)
    b = (u[idx] / f32(dt)).astype(f32)
    scale = (f32(sigma) * delta(sample(phi, tuple(I[a].astype(f32) + off[a] for a in range(dim))),
                                dx)).astype(f32)
    grad = []
    for k in range(dim):
        pp = tuple(I[a].astype(f32) + f32(0.5 if a == k else 0.0) for a in range(dim))
        pm = tuple(I[a].astype(f32) - f32(0.5 if a == k else 0.0) for a in range(dim))
        g = ((sample(u, pp) - sample(u, pm)) / f32(dx)).astype(f32)
        grad.append(g)
    nk = [n[..., a] for a in range(dim)]
    Du = (grad[0] * nk[0]).astype(f32)
    for a in range(1, dim):
        Du = (Du + grad[a] * nk[a]).astype(f32)
    dx2 = f32(dx ** 2)
    D2 = [[None] * dim for _ in range(dim)]
    for k1 in range(dim):
        for k2 in range(dim):
            def p(c1, c2):
                return tuple((I[a].astype(f32) + f32(c1 * (a == k1)) + f32(c2 * (a == k2))).astype(f32)
                             if k1 != k2 else
                             (I[a].astype(f32) + f32((c1 + c2) * (a == k1))).astype(f32)
                             for a in range(dim))
            v00 = sample(u, p(-0.5, -0.5))
            v10 = sample(u, p(0.5, -0.5))
            v01 = sample(u, p(-0.5, 0.5))
            v11 = sample(u, p(0.5, 0.5))
            D2[k1][k2] = ((((v11 + v00) - v10) - v01) / dx2).astype(f32)
    row = []
    for j in range(dim):
        acc = (nk[0] * D2[0][j]).astype(f32)
        for i in range(1, dim):
            acc = (acc + nk[i] * D2[i][j]).astype(f32)
        row.append(acc)
    D2n = (row[0] * nk[0]).astype(f32)
    for j in range(1, dim):
        D2n = (D2n + row[j] * nk[j]).astype(f32)
    inner = (D2n + (kappa * Du).astype(f32)).astype(f32)
    tot = ((kappa * nk[d]).astype(f32) + (f32(dt) * inner).astype(f32)).astype(f32)
    b = (b - (scale * tot).astype(f32)).astype(f32)
    return b


def surface_build_A(res, level, d, phi, dt, dx, sigma=7.28e-2):
    """build_A_kernel (surface_tension.py:71-89) at one level -> (Adiag, Ax).

    The bound ``I[k] + 1 < res[k]`` uses the fine ``res`` at every level
    (surface_tension.py:82).
    """
    dim = len(res)
    shape = tuple(n >> level for n in res)
    I = _grid(shape)
    # ``dx`` is a kernel-scope f32 variable (surface_tension.py:73): the scale is
    # f32(sigma * dt) / (f32(dx_l) * f32(dx_l)) (surface_tension.py:78)
    dxl = f32(dx * (2 ** level))
    scale = (f32(sigma * dt) / (dxl * dxl)).astype(f32)
    lvl = f32(2 ** level)
    off = [f32(0.0 if a == d else 0.5) for a in range(dim)]
    Adiag = np.full(shape, f32(1 / dt), f32)
    Ax = np.zeros(shape + (dim,), f32)
    for k in range(dim):
        up = tuple((((I[a].astype(f32) + f32(0.5 if a == k else 0.0)) + off[a]) * lvl).astype(f32)
                   for a in range(dim))
        w_up = (scale * delta(sample(phi, up), dx)).astype(f32)
        m_up = I[k] + 1 < res[k]
        Ax[..., k] = np.where(m_up, (f32(-scale) * delta(sample(phi, up), dx)).astype(f32),
                              Ax[..., k])
        Adiag = np.where(m_up, (Adiag + w_up).astype(f32), Adiag)
        dn = tuple((((I[a].astype(f32) - f32(0.5 if a == k else 0.0)) + off[a]) * lvl).astype(f32)
                   for a in range(dim))
        w_dn = (scale * delta(sample(phi, dn), dx)).astype(f32)
        m_dn = I[k] - 1 >= 0
        Adiag = np.where(m_dn, (Adiag + w_dn).astype(f32), Adiag)
    return Adiag, Ax


# ----------------------------------------------------------------------------
# Hierarchy assembly (mgpcg.py:79-94)
# ----------------------------------------------------------------------------
def hierarchy_from_cell_type(cell_type, L, build_A):
    """reinitialize (mgpcg.py:79-87): coarsen cell types, build_A per level."""
    from . import core

    gt = [np.ascontiguousarray(cell_type, np.int8)]
    for _ in range(1, L):
        gt.append(core.coarsen(gt[-1]))
    Ad, Ax = [], []
    for l in range(L):
        a, x = build_A(gt[l], l)
        Ad.append(np.ascontiguousarray(a, f32))
        Ax.append(np.ascontiguousarray(x, f32))
    return {"gt": gt, "Adiag": Ad, "Ax": Ax}


def hierarchy_full(res, L, build_A):
    """full_reinitialize (mgpcg.py:89-94): every level all FLUID."""
    gt, Ad, Ax = [], [], []
    for l in range(L):
        shape = tuple(n >> l for n in res)
        gt.append(np.zeros(shape, np.int8))
        a, x = build_A(l)
        Ad.append(np.ascontiguousarray(a, f32))
        Ax.append(np.ascontiguousarray(x, f32))
    return {"gt": gt, "Adiag": Ad, "Ax": Ax}
