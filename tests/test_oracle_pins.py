"""Pins of the oracle by what the mathematics fixes, each with a planted slip.

The reference ships no tests, fixtures or golden vectors, and Taichi cannot
be installed here, so the oracle (oracle/strategies.py, oracle/mgpcg_oracle.c)
is pinned by closed forms, invariants and independent assemblies.  Every
check below is paired with a plausible mistake planted in a COPY of the
oracle (the numpy function's source with one substitution, or the C file
with one substitution, compiled into a temporary library); the same check
must reject the planted copy.  The substitution must be present in the
oracle's text, so a test also fails when the oracle is edited under it.

  sample            exact on multilinear data, clamped to the edge values
  theta / delta     theta(0) = 1/2, symmetry, d theta / d phi = delta,
                    integral of delta = 1, delta(0) = 1 / eps, eps = 5 dx
  potential A       uniform chi: -(1 + chi) / dx_l times the graph Laplacian,
                    Neumann at level 0, Dirichlet on the upper faces of the
                    coarse levels (the fine-res domain test): cosine eigenpairs
  potential A       chi varying along one axis: face values of the linear
                    interpolation (the face-midpoint average at level 0, the
                    fine cell value at 2^l I + 2^(l-1) e_k on coarse levels)
  potential b       piecewise-constant chi, uniform H: the chi jump times H
  surface A         |phi| >= eps everywhere: A = I / dt; a flat interface:
                    couplings -sigma dt / dx_l^2 delta(phi at the face),
                    A - I / dt with zero row sums
  surface b         flat interface, quadratic velocity: the closed form
  calc_n, kappa     plane: n = e_y, kappa = 0; sphere: kappa = (d-1) / rho
  calc_volume       half-space exactly; sphere: the smoothed volume within O(dx^2)
  PI controller     three steps computed by hand
  magnetic b        uniform H: pressure b with p0 + k mu0 |H|^2 / 2;
                    linear H: the AIR cell-centre value of |H|^2
  apply_magnetic    linear H: the AIR cell-centre value on fluid-air faces
  restrict          R (r - A z) with z != 0 against a scipy assembly
  solve (FP64)      a textbook PCG with the same preconditioner: the same
                    iterates and iteration count; PLAIN FP32 within 1e-4
  C oracle plants   the pins of test_oracle_core.py (coarsening, neighbour
                    sum, colouring, prolongation, V-cycle order) reject a
                    planted copy
"""
import contextlib
import ctypes
import inspect
import math
import os
import subprocess
import textwrap

import numpy as np
import pytest

from helpers import AIR, FLUID, SOLID, assemble_matrix, random_hierarchy, rel_l2
from oracle import core
from oracle import strategies as S

f32 = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_C = os.path.join(ROOT, "oracle", "mgpcg_oracle.c")


# ----------------------------------------------------------------------------
# planting helpers
# ----------------------------------------------------------------------------
def planted(func, old, new):
    """A copy of the oracle function ``func`` with ``old`` replaced by ``new``
    in its source, compiled in the oracle module's namespace."""
    src = textwrap.dedent(inspect.getsource(func))
    assert old in src, (func.__name__, old)
    ns = dict(func.__globals__)
    exec(compile(src.replace(old, new), f"<planted {func.__name__}>", "exec"), ns)
    return ns[func.__name__]


_C_CACHE = {}


def planted_c(tmp_path_factory, subs):
    """The C oracle with each (old, new) substitution applied (every
    occurrence), compiled into a temporary library."""
    key = tuple(subs)
    if key in _C_CACHE:
        return _C_CACHE[key]
    src = open(ORACLE_C).read()
    for old, new in subs:
        assert old in src, old
        src = src.replace(old, new)
    d = tmp_path_factory.mktemp("planted_oracle")
    c = d / "oracle.c"
    c.write_text(src)
    so = d / "oracle.so"
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                    "-std=c11", "-w", str(c), "-o", str(so), "-lm"], check=True)
    lib = ctypes.CDLL(str(so))
    core._declare(lib)
    _C_CACHE[key] = lib
    return lib


@contextlib.contextmanager
def oracle_lib(lib):
    """Run the oracle's ctypes wrappers on ``lib`` (None: the real oracle)."""
    if lib is None:
        yield
        return
    old = core._lib
    core._lib = lib
    try:
        yield
    finally:
        core._lib = old


def _grid(shape):
    return np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")


# ----------------------------------------------------------------------------
# utils.sample (utils.py:12-32)
# ----------------------------------------------------------------------------
def _check_sample(fn, dim):
    rng = np.random.default_rng(40 + dim)
    shape = (7, 9, 6)[:dim]
    coef = rng.uniform(-1, 1, 2 ** dim)

    def f(x):
        v = np.zeros_like(x[0], np.float64)
        for m, c in enumerate(coef):  # every product of a subset of the axes
            t = np.full_like(x[0], c, np.float64)
            for a in range(dim):
                if (m >> a) & 1:
                    t = t * x[a]
            v = v + t
        return v

    G = [g.astype(np.float64) for g in _grid(shape)]
    data = f(G).astype(f32)
    npts = 400
    pos = [rng.uniform(-1.5, n + 0.5, npts).astype(f32) for n in shape]
    got = fn(data, tuple(pos)).astype(np.float64)
    clipped = [np.clip(p.astype(np.float64), 0, n - 1) for p, n in zip(pos, shape)]
    exp = f(clipped)
    scale = np.abs(data).max()
    return float(np.max(np.abs(got - exp)) / scale)


@pytest.mark.parametrize("dim", [2, 3])
def test_sample_exact_on_multilinear_data_and_clamped(dim):
    assert _check_sample(S.sample, dim) < 1e-5


@pytest.mark.parametrize("old,new", [
    ("ip = np.clip(i + 1, 0, n - 1)", "ip = np.clip(i, 0, n - 1)"),
    ("d[i, j] * (one - s) + d[ip, j] * s", "d[i, j] * s + d[ip, j] * (one - s)"),
    ("d[i, j, k] * (one - s) + d[ip, j, k] * s", "d[i, j, k] * s + d[ip, j, k] * (one - s)"),
    ("s = np.clip(p - i.astype(f32), f32(0.0), f32(1.0))", "s = p - i.astype(f32)"),
])
def test_sample_check_rejects_planted(old, new):
    bad = planted(S.sample, old, new)
    assert max(_check_sample(bad, 2), _check_sample(bad, 3)) > 1e-3


# ----------------------------------------------------------------------------
# LevelSet.theta / delta (level_set.py:22-37)
# ----------------------------------------------------------------------------
def _check_theta_delta(theta, delta, dx=1.0 / 64):
    eps = 5 * dx  # level_set.py:22
    errs = []
    phi = np.linspace(-1.5 * eps, 1.5 * eps, 2001)
    th = theta(phi.astype(f32), dx).astype(np.float64)
    de = delta(phi.astype(f32), dx).astype(np.float64)
    errs.append(abs(theta(np.array([0.0], f32), dx)[0] - 0.5))
    errs.append(abs(theta(np.array([-eps], f32), dx)[0] - 0.0))
    errs.append(abs(theta(np.array([eps], f32), dx)[0] - 1.0))
    errs.append(np.max(np.abs(th + th[::-1] - 1.0)))  # theta(-phi) = 1 - theta(phi)
    errs.append(abs(np.trapezoid(de, phi) - 1.0))
    errs.append(abs(delta(np.array([0.0], f32), dx)[0] * eps - 1.0))
    errs.append(abs(delta(np.array([eps], f32), dx)[0]) * eps)
    dth = np.gradient(th, phi)
    inner = np.abs(phi) < 0.98 * eps
    errs.append(np.max(np.abs(dth[inner] - de[inner])) * eps)
    return float(max(errs))


def test_theta_delta_closed_forms():
    assert _check_theta_delta(S.theta, S.delta) < 1e-4


@pytest.mark.parametrize("which,old,new", [
    ("theta", "f32(1 / (2 * math.pi)) * np.sin(", "f32(1 / math.pi) * np.sin("),
    ("theta", "(f32(0.5) + phi / f32(2 * eps))", "(f32(0.5) + phi / f32(eps))"),
    ("delta", "/ f32(2 * eps)", "/ f32(eps)"),
    ("delta", "eps = 5 * dx", "eps = 4 * dx"),
    ("delta", "np.cos(f32(math.pi) * phi / f32(eps))", "np.cos(f32(math.pi) * phi / f32(2 * eps))"),
])
def test_theta_delta_check_rejects_planted(which, old, new):
    th, de = S.theta, S.delta
    if which == "theta":
        th = planted(S.theta, old, new)
    else:
        de = planted(S.delta, old, new)
    assert _check_theta_delta(th, de) > 1e-3


# ----------------------------------------------------------------------------
# PotentialFunctionStrategy.build_A (potential_function.py:40-55)
# ----------------------------------------------------------------------------
def _cos_mode(shape, modes, coarse):
    """Eigenvector and Laplacian eigenvalue (of L = degree - adjacency) of the
    graph Laplacian with Neumann lower faces and, when ``coarse``, Dirichlet
    upper faces (ghost value 0); Neumann on both sides otherwise."""
    G = _grid(shape)
    v = np.ones(shape)
    lam = 0.0
    for a, (n, m) in enumerate(zip(shape, modes)):
        th = math.pi * (m + 0.5) / (n + 0.5) if coarse else math.pi * m / n
        v = v * np.cos(th * (G[a] + 0.5))
        lam += 2.0 - 2.0 * math.cos(th)
    return v, lam


def _check_potential_uniform(build_A, res, dx, chi0, levels):
    """Worst relative eigen-residual over the levels (rows away from the
    reference cell and its neighbours), plus the reference row."""
    worst = 0.0
    dim = len(res)
    chi = np.full(res, f32(chi0), f32)
    for l in levels:
        Ad, Ax = build_A(res, l, dx, chi)
        shape = Ad.shape
        gt = np.zeros(shape, np.int8)
        A = assemble_matrix(gt, Ad.astype(np.float64), Ax.astype(np.float64), fluid_only=False)
        c = (1.0 + chi0) / (dx * 2 ** l)
        modes = (1, 2, 1)[:dim]
        v, lam = _cos_mode(shape, modes, coarse=l > 0)
        Av = (A @ v.ravel()).reshape(shape)
        keep = np.ones(shape, bool)
        keep[(0,) * dim] = False
        for k in range(dim):
            idx = [0] * dim
            idx[k] = 1
            keep[tuple(idx)] = False
        err = np.abs(Av + c * lam * v)[keep].max() / (c * lam * np.abs(v).max())
        worst = max(worst, err)
        # reference cell (psi = 0 there): Adiag = 1 / dx_l, no couplings
        worst = max(worst, abs(Ad[(0,) * dim] * dx * 2 ** l - 1.0),
                    float(np.abs(Ax[(0,) * dim]).max()))
    return worst


@pytest.mark.parametrize("res", [(16, 12), (16, 12, 8)])
def test_potential_A_uniform_chi_is_scaled_laplacian(res):
    assert _check_potential_uniform(S.potential_build_A, res, 1.0 / 64, 2.0, [0, 1, 2]) < 1e-5


@pytest.mark.parametrize("old,new", [
    ("dxl = dx * (2 ** level)", "dxl = dx"),
    ("m = dom & _in_domain(nbc, res)", "m = dom & _in_domain(nbc, shape)"),
    ("Adiag = np.where(m, (Adiag - val).astype(f32), Adiag)",
     "Adiag = np.where(m, (Adiag + val).astype(f32), Adiag)"),
    ("Adiag = np.where(origin, inv, Adiag)", "Adiag = Adiag"),
])
def test_potential_uniform_check_rejects_planted(old, new):
    bad = planted(S.potential_build_A, old, new)
    assert _check_potential_uniform(bad, (16, 12, 8), 1.0 / 64, 2.0, [0, 1, 2]) > 1e-3


def _check_potential_faces(build_A, res, dx):
    """chi varies along axis 0 only: the coupling of the face between I and
    I + e_0 is 1 + chi interpolated linearly at its midpoint (level 0), or
    the fine cell value at 2^l I_0 + 2^(l-1) (level l > 0); along the other
    axes chi is constant, so their couplings carry the cell's own value."""
    rng = np.random.default_rng(50)
    g = rng.uniform(0.0, 3.0, res[0]).astype(f32)
    chi = np.broadcast_to(g.reshape((-1,) + (1,) * (len(res) - 1)), res).astype(f32).copy()
    worst = 0.0
    for l in (0, 1, 2):
        Ad, Ax = build_A(res, l, dx, chi)
        n0 = Ad.shape[0]
        dxl = dx * 2 ** l
        i = np.arange(n0 - 1)
        if l == 0:
            face = 0.5 * (g[i].astype(np.float64) + g[i + 1])
        else:
            face = g[(2 ** l) * i + 2 ** (l - 1)].astype(np.float64)
        exp = (1.0 + face) / dxl
        got = Ax[(slice(0, n0 - 1),) + (slice(1, None),) * (len(res) - 1) + (0,)]
        got = got.reshape(n0 - 1, -1).astype(np.float64)
        worst = max(worst, float(np.max(np.abs(got - exp[:, None]) / exp[:, None])))
        # axis 1: the cell's own chi (coarse levels: the fine cell 2^l I_0)
        own = g[(2 ** l) * np.arange(n0)].astype(np.float64)
        sl = (slice(1, n0),) + (slice(0, Ad.shape[1] - 1),) + (slice(None),) * (len(res) - 2)
        got1 = Ax[sl + (1,)].reshape(n0 - 1, -1).astype(np.float64)
        exp1 = (1.0 + own[1:]) / dxl
        worst = max(worst, float(np.max(np.abs(got1 - exp1[:, None]) / exp1[:, None])))
    return worst


def test_potential_A_face_sampling():
    assert _check_potential_faces(S.potential_build_A, (32, 8, 8), 1.0 / 32) < 1e-6


@pytest.mark.parametrize("old,new", [
    ("f32(0.5 * s if a == k else 0.0)", "f32(1.0 * s if a == k else 0.0)"),
    ("* scale).astype(f32)", ").astype(f32)"),
])
def test_potential_faces_check_rejects_planted(old, new):
    bad = planted(S.potential_build_A, old, new)
    assert _check_potential_faces(bad, (32, 8, 8), 1.0 / 32) > 1e-3


# ----------------------------------------------------------------------------
# PotentialFunctionStrategy.build_b (potential_function.py:19-34)
# ----------------------------------------------------------------------------
def _check_potential_b(build_b, res):
    """chi = chi_lo below row j0 of axis 1 and chi_hi from it on, H uniform:
    inside the domain b = H_1 (chi_{j+1} - chi_j), the jump times H, nonzero
    on row j0 - 1 only; b = 0 at the reference cell."""
    dim = len(res)
    j0 = res[1] // 2
    lo, hi = f32(0.5), f32(2.0)
    G = _grid(res)
    chi = np.where(G[1] < j0, lo, hi).astype(f32)
    h = [f32(0.7), f32(-1.3), f32(0.4)][:dim]
    H = []
    for k in range(dim):
        shp = list(res)
        shp[k] += 1
        H.append(np.full(shp, h[k], f32))
    b = build_b(list(res), chi, H).astype(np.float64)
    inner = np.ones(res, bool)
    for a in range(dim):
        inner &= (G[a] > 0) & (G[a] < res[a] - 1)
    exp = np.where(G[1] == j0 - 1, float(h[1]) * (float(hi) - float(lo)), 0.0)
    err = float(np.max(np.abs(b - exp)[inner]))
    err = max(err, abs(b[(0,) * dim]))
    return err


@pytest.mark.parametrize("res", [(12, 10), (8, 10, 6)])
def test_potential_b_is_chi_jump_times_H(res):
    assert _check_potential_b(S.potential_build_b, res) < 1e-6


@pytest.mark.parametrize("old,new", [
    ("pos_up = tuple(up[a].astype(f32) + off[a]", "pos_up = tuple(I[a].astype(f32) + off[a]"),
    ("b = np.where(m_dn, (b - term).astype(f32), b)", "b = np.where(m_dn, (b + term).astype(f32), b)"),
    ("off = [f32(0.0 if a == k else 0.5) for a in range(dim)]",
     "off = [f32(0.5 if a == k else 0.0) for a in range(dim)]"),
])
def test_potential_b_check_rejects_planted(old, new):
    bad = planted(S.potential_build_b, old, new)
    assert _check_potential_b(bad, (8, 10, 6)) > 1e-3


# ----------------------------------------------------------------------------
# SurfaceTensionStrategy.build_A (surface_tension.py:71-89)
# ----------------------------------------------------------------------------
def _delta_exact(phi, dx):
    eps = 5 * dx
    return np.where(np.abs(phi) < eps, (1 + np.cos(np.pi * phi / eps)) / (2 * eps), 0.0)


def _check_surface_A(build_A, res, d, dt=1e-2, dx=1.0 / 32, sigma=7.28e-2):
    """(1) |phi| >= eps everywhere: A = I / dt exactly.  (2) a flat interface
    phi = y - y0: each coupling is -sigma dt / dx_l^2 delta(phi) at the face
    (I + e_k / 2 + offset_d) 2^l, and A - I / dt has zero row sums on the
    interior (each lower face weight equals the neighbour's coupling)."""
    dim = len(res)
    worst = 0.0
    far = np.full(res, f32(1.0), f32)
    for l in (0, 1):
        Ad, Ax = build_A(list(res), l, d, far, dt, dx, sigma)
        worst = max(worst, float(np.max(np.abs(Ad - f32(1.0 / dt)))) * dt,
                    float(np.max(np.abs(Ax))))
    y0 = 0.47
    Gf = _grid(res)
    phi = ((Gf[1] + 0.5) * dx - y0).astype(f32)
    off = [0.0 if a == d else 0.5 for a in range(dim)]
    for l in (0, 1):
        Ad, Ax = build_A(list(res), l, d, phi, dt, dx, sigma)
        shape = Ad.shape
        G = _grid(shape)
        dxl = dx * 2 ** l
        scale = sigma * dt / dxl ** 2
        for k in range(dim):
            py = (G[1] + (0.5 if k == 1 else 0.0) + off[1]) * 2 ** l
            ok = (py >= 0) & (py <= res[1] - 1) & (G[k] + 1 < shape[k])
            exp = -scale * _delta_exact((py + 0.5) * dx - y0, dx)
            worst = max(worst, float(np.max(np.abs(Ax[..., k] - exp)[ok])) / scale * dx)
        if l == 0:
            A = assemble_matrix(np.zeros(shape, np.int8), Ad.astype(np.float64),
                                Ax.astype(np.float64), fluid_only=False)
            rows = (A @ np.ones(A.shape[0])).reshape(shape) - 1.0 / dt
            inner = np.ones(shape, bool)
            for a in range(dim):
                inner &= (G[a] > 0) & (G[a] < shape[a] - 1)
            worst = max(worst, float(np.max(np.abs(rows[inner]))) / scale * dx)
    return worst


@pytest.mark.parametrize("res,d", [((16, 32), 0), ((16, 32), 1), ((8, 32, 8), 2)])
def test_surface_A_closed_forms(res, d):
    assert _check_surface_A(S.surface_build_A, res, d) < 1e-4


@pytest.mark.parametrize("old,new", [
    ("Adiag = np.full(shape, f32(1 / dt), f32)", "Adiag = np.full(shape, f32(dt), f32)"),
    ("dn = tuple((((I[a].astype(f32) - f32(0.5 if a == k else 0.0))",
     "dn = tuple((((I[a].astype(f32) + f32(0.5 if a == k else 0.0))"),
    ("+ off[a]) * lvl).astype(f32)\n                   for a in range(dim))\n        w_up",
     "+ off[a])).astype(f32)\n                   for a in range(dim))\n        w_up"),
    ("off = [f32(0.0 if a == d else 0.5) for a in range(dim)]",
     "off = [f32(0.0) for a in range(dim)]"),
])
def test_surface_A_check_rejects_planted(old, new):
    bad = planted(S.surface_build_A, old, new)
    assert max(_check_surface_A(bad, (16, 32), 0), _check_surface_A(bad, (8, 32, 8), 2)) > 1e-3


# ----------------------------------------------------------------------------
# SurfaceTensionStrategy.build_b (surface_tension.py:23-66)
# ----------------------------------------------------------------------------
def _check_surface_b(build_b, res, d, K, dt=1e-2, dx=1.0 / 32, sigma=7.28e-2):
    """Flat interface phi = y - y0 (n = e_y), curvature K, velocity component
    d = a j^2 (j the index along axis 1): D2u/Dn2 = 2a / dx^2 and
    Du/Dn = 2 a j / dx, so on rows 1..n-2 of axis 1
      b = a j^2 / dt - sigma delta(phi) (K n_d + dt (2a / dx^2 + K 2 a j / dx))
    with phi taken at I + offset_d."""
    dim = len(res)
    a = 0.37
    ushape = list(res)
    ushape[d] += 1
    Gu = _grid(ushape)
    u = (a * Gu[1].astype(np.float64) ** 2).astype(f32)
    y0 = 0.47
    Gf = _grid(res)
    phi = ((Gf[1] + 0.5) * dx - y0).astype(f32)
    n = np.zeros(tuple(res) + (dim,), f32)
    n[..., 1] = 1.0
    kappa = np.full(res, f32(K), f32)
    b = build_b(u, phi, kappa, n, d, dt, dx, sigma).astype(np.float64)
    j = Gf[1].astype(np.float64)
    off1 = 0.0 if d == 1 else 0.5
    de = _delta_exact((j + off1 + 0.5) * dx - y0, dx)
    nd = 1.0 if d == 1 else 0.0
    exp = a * j ** 2 / dt - sigma * de * (K * nd + dt * (2 * a / dx ** 2 + K * 2 * a * j / dx))
    ok = (Gf[1] >= 1) & (Gf[1] <= res[1] - 2)
    mag = np.abs(sigma * de * dt * 2 * a / dx ** 2).max()
    # u / dt is large beside the surface term: allow its FP32 rounding
    slack = 4 * np.finfo(f32).eps * np.abs(a * j ** 2 / dt)
    return float(np.max(np.maximum(np.abs(b - exp) - slack, 0.0)[ok]) / mag)


@pytest.mark.parametrize("res,d,K", [((16, 32), 0, 0.0), ((16, 32), 1, 3.0),
                                     ((8, 32, 8), 2, -2.0), ((8, 32, 8), 1, 5.0)])
def test_surface_b_closed_form(res, d, K):
    assert _check_surface_b(S.surface_build_b, res, d, K) < 1e-4


@pytest.mark.parametrize("old,new", [
    ("D2[k1][k2] = ((((v11 + v00) - v10) - v01) / dx2)",
     "D2[k1][k2] = ((((v11 - v00) - v10) + v01) / dx2)"),
    ("tot = ((kappa * nk[d]).astype(f32) + (f32(dt) * inner).astype(f32))",
     "tot = ((kappa * nk[d]).astype(f32) + inner)"),
    ("inner = (D2n + (kappa * Du).astype(f32))", "inner = D2n"),
    ("delta(sample(phi, tuple(I[a].astype(f32) + off[a] for a in range(dim)))",
     "delta(sample(phi, tuple(I[a].astype(f32) for a in range(dim)))"),
])
def test_surface_b_check_rejects_planted(old, new):
    bad = planted(S.surface_build_b, old, new)
    assert max(_check_surface_b(bad, (16, 32), 1, 3.0), _check_surface_b(bad, (8, 32, 8), 2, -2.0),
               _check_surface_b(bad, (16, 32), 0, 0.0)) > 1e-3


# ----------------------------------------------------------------------------
# SurfaceTension.calc_n / calc_kappa (surface_tension.py:115-136)
# ----------------------------------------------------------------------------
def _check_n_kappa(calc_n, calc_kappa, dim):
    """Plane phi = y - y0: n = e_y and kappa = 0.  Sphere phi = |x - c| - r:
    kappa = (d - 1) / rho at the sample position I + offset_d, within 0.5%
    on the cells within 2 dx of the interface."""
    n_side = 48 if dim == 3 else 96
    dx = 1.0 / n_side
    res = (n_side,) * dim
    G = _grid(res)
    worst = 0.0
    plane = ((G[1] + 0.5) * dx - 0.43).astype(f32)
    for d in range(dim):
        nn = calc_n(plane, d, dx)
        ey = np.zeros(dim)
        ey[1] = 1.0
        # rows whose samples stay inside the grid (the last row's upper sample
        # is clamped when the offset moves it up half a cell)
        rows = (G[1] >= 1) & (G[1] <= n_side - 3)
        worst = max(worst, float(np.max(np.abs(nn - ey)[rows])))
        worst = max(worst, float(np.max(np.abs(calc_kappa(nn, dx))[rows])) * dx)
    r = 0.3
    rho = np.sqrt(sum(((G[a] + 0.5) * dx - 0.5) ** 2 for a in range(dim)))
    phi = (rho - r).astype(f32)
    for d in range(dim):
        kap = calc_kappa(calc_n(phi, d, dx), dx).astype(np.float64)
        off = [0.0 if a == d else 0.5 for a in range(dim)]
        rp = np.sqrt(sum(((G[a] + off[a] + 0.5) * dx - 0.5) ** 2 for a in range(dim)))
        band = np.abs(rp - r) < 2 * dx
        ex = (dim - 1) / rp
        worst = max(worst, float(np.max(np.abs(kap - ex)[band] / ex[band])) * 2e-4 / 5e-3)
    return worst


@pytest.mark.parametrize("dim", [2, 3])
def test_calc_n_kappa_plane_and_sphere(dim):
    assert _check_n_kappa(S.calc_n, S.calc_kappa, dim) < 2e-4


@pytest.mark.parametrize("which,old,new", [
    ("n", "pp = tuple((I[a].astype(f32) + f32(0.5 if a == k else 0.0)) + off[a]",
     "pp = tuple((I[a].astype(f32) + f32(0.5 if a == k else 0.0))"),
    ("n", "n = np.where((norm > 0)[..., None], (n / safe[..., None]).astype(f32), n)", "n = n"),
    ("kappa", "kappa = (kappa + (diff / f32(dx)).astype(f32)).astype(f32)",
     "kappa = (kappa + diff).astype(f32)"),
    ("kappa", "diff = (sample(n[..., k], pp) - sample(n[..., k], pm))",
     "diff = (sample(n[..., k], pp) - sample(n[..., k], tuple(I[a].astype(f32) for a in range(dim))))"),
])
def test_n_kappa_check_rejects_planted(which, old, new):
    cn, ck = S.calc_n, S.calc_kappa
    if which == "n":
        cn = planted(S.calc_n, old, new)
    else:
        ck = planted(S.calc_kappa, old, new)
    assert max(_check_n_kappa(cn, ck, 2), _check_n_kappa(cn, ck, 3)) > 1e-3


# ----------------------------------------------------------------------------
# PressureProjectWithVolumeControlStrategy (volume_control.py:13-51)
# ----------------------------------------------------------------------------
def _theta_exact(p, eps):
    if p <= -eps:
        return 0.0
    if p >= eps:
        return 1.0
    return 0.5 + p / (2 * eps) + math.sin(math.pi * p / eps) / (2 * math.pi)


def _check_volume(calc_volume, dim):
    """Half-space below y0: y0 / dx cells per column (the smoothed step is
    odd about the interface, so a plane loses nothing).  A sphere of radius
    r: the radial integral of theta(r - rho) (the smoothed volume; the
    smoothing itself adds O(eps^2) to 4/3 pi r^3), within 1% in 3D and 0.1%
    in 2D (the corner values of the non-linear phi are averages of cell
    values, an O(dx^2) bias).  Returns the larger error over its bound."""
    from scipy.integrate import quad

    n_side = 48 if dim == 3 else 96
    dx = 1.0 / n_side
    eps = 5 * dx
    res = (n_side,) * dim
    G = _grid(res)
    y0 = 0.4
    half = ((G[1] + 0.5) * dx - y0).astype(f32)
    cols = n_side ** (dim - 1)
    e1 = abs(calc_volume(half, dx) - cols * y0 / dx) / (cols * y0 / dx)
    r = 0.3
    rho = np.sqrt(sum(((G[a] + 0.5) * dx - 0.5) ** 2 for a in range(dim)))
    ball = (rho - r).astype(f32)
    shell = (lambda q: 4 * math.pi * q * q) if dim == 3 else (lambda q: 2 * math.pi * q)
    smooth, _ = quad(lambda q: _theta_exact(r - q, eps) * shell(q), 0.0, r + 2 * eps,
                     points=[r - eps, r, r + eps], limit=200)
    smooth /= dx ** dim
    e2 = abs(calc_volume(ball, dx) - smooth) / smooth
    return max(e1 / 1e-6, e2 / (1e-2 if dim == 3 else 1e-3))


@pytest.mark.parametrize("dim", [2, 3])
def test_calc_volume_half_space_and_sphere(dim):
    assert _check_volume(S.calc_volume, dim) < 1.0


@pytest.mark.parametrize("old,new", [
    ("tot = (tot + theta(-sample(phi, pos), dx))", "tot = (tot + theta(sample(phi, pos), dx))"),
    ("(I[a] + off[a]).astype(f32) - f32(0.5)", "(I[a] + off[a]).astype(f32) + f32(0.5)"),
    ("cell = (tot / f32(2 ** dim))", "cell = (tot / f32(2 ** (dim - 1)))"),
])
def test_volume_check_rejects_planted(old, new):
    bad = planted(S.calc_volume, old, new)
    assert max(_check_volume(bad, 2), _check_volume(bad, 3)) > 10.0


def _pi_steps(ctrl_cls):
    c = ctrl_cls(dt=0.01)
    return [c.update(v) for v in (100.0, 101.0, 99.5)]


def test_volume_controller_three_steps_by_hand():
    """dt = 0.01: k_p = 2.3 / (25 dt) = 9.2, k_i = (k_p / (2 zeta))^2 = 5.29.
    step 1 records vol_0 = 100 and returns nothing; step 2 (vol 101):
    x = 0.01, y = 1e-4, c = (-0.092 - 5.29e-4) / 1.01; step 3 (vol 99.5):
    x = -0.005, y = 5e-5, c = (0.046 - 2.645e-4) / 0.995."""
    got = _pi_steps(S.VolumeController)
    assert got[0] is None
    assert got[1] == pytest.approx(-0.092529 / 1.01, rel=1e-12)
    assert got[2] == pytest.approx(0.0457355 / 0.995, rel=1e-12)


@pytest.mark.parametrize("old,new", [
    ("self.k_i = (self.k_p / (2 * zeta)) ** 2", "self.k_i = (self.k_p / zeta) ** 2"),
    ("self.y += x * self.dt", "self.y += x"),
    ("return (-self.k_p * x - self.k_i * self.y) / (x + 1)",
     "return (-self.k_p * x - self.k_i * self.y)"),
])
def test_volume_controller_check_rejects_planted(old, new):
    src = textwrap.dedent(inspect.getsource(S.VolumeController))
    assert old in src
    ns = dict(S.__dict__)
    exec(compile(src.replace(old, new), "<planted VolumeController>", "exec"), ns)
    got = _pi_steps(ns["VolumeController"])
    assert abs(got[1] - (-0.092529 / 1.01)) > 1e-6 or abs(got[2] - 0.0457355 / 0.995) > 1e-6


# ----------------------------------------------------------------------------
# MagneticForceStrategy.build_b (magnetic_force.py:17-39) and
# apply_magnetic_force (ferrofluid_simulator.py:114-137)
# ----------------------------------------------------------------------------
def _magnetic_scene(res, seed):
    rng = np.random.default_rng(seed)
    dim = len(res)
    u = rng.uniform(size=res)
    ct = np.where(u < 0.55, FLUID, np.where(u < 0.85, AIR, SOLID)).astype(np.int8)
    vel = []
    for k in range(dim):
        shp = list(res)
        shp[k] += 1
        vel.append(rng.uniform(-1, 1, shp).astype(f32))
    return ct, vel


def _face_H(res, h):
    """H_k on the faces of axis k (shape res + e_k), constant h_k."""
    out = []
    for k in range(len(res)):
        shp = list(res)
        shp[k] += 1
        out.append(np.full(shp, f32(h[k]), f32))
    return out


def _linear_Hx(res, a):
    """H_0 = a * (face position along axis 0), the other components 0."""
    H = _face_H(res, [0.0] * len(res))
    G = _grid(H[0].shape)
    H[0] = (a * G[0].astype(np.float64)).astype(f32)
    return H


MU0, KK = 1.25663706e-6, 2.0e5


def _check_magnetic_b(build_b, res):
    sA, sb, p0 = 0.8, 1.3, 0.25
    ct, vel = _magnetic_scene(res, 60)
    h = [0.6, -0.9, 0.3][:len(res)]
    hh = sum(v * v for v in h)
    b = build_b(ct, vel, _face_H(res, h), sA, sb, p0, MU0, KK).astype(np.float64)
    bp = S.pressure_build_b(ct, vel, sA, sb, p0 + 0.5 * KK * MU0 * hh,
                            ghost_fluid_method=False).astype(np.float64)
    e1 = np.max(np.abs(b - bp)) / np.max(np.abs(bp))
    # linear H_0 = a x: |H|^2 at the AIR cell J's centre is (a (J_0 + 1/2))^2
    a = 3.0
    b2 = build_b(ct, vel, _linear_Hx(res, a), sA, sb, 0.0, MU0, KK).astype(np.float64)
    b0 = S.pressure_build_b(ct, vel, sA, sb, 0.0, ghost_fluid_method=False).astype(np.float64)
    exp = np.zeros(res)
    G = _grid(res)
    for k in range(len(res)):
        for s in (-1, 1):
            nb = np.full(res, SOLID, np.int8)
            src = [slice(None)] * len(res)
            dst = [slice(None)] * len(res)
            if s > 0:
                src[k], dst[k] = slice(1, None), slice(None, -1)
            else:
                src[k], dst[k] = slice(None, -1), slice(1, None)
            nb[tuple(dst)] = ct[tuple(src)]
            Jx = G[0] + (s if k == 0 else 0)
            term = sA * 0.5 * KK * MU0 * (a * (Jx + 0.5)) ** 2
            exp += np.where((ct == FLUID) & (nb == AIR), term, 0.0)
    e2 = np.max(np.abs((b2 - b0) - exp)) / np.max(np.abs(exp))
    return float(max(e1, e2))


@pytest.mark.parametrize("res", [(10, 9), (6, 7, 5)])
def test_magnetic_b_closed_forms(res):
    assert _check_magnetic_b(S.magnetic_build_b, res) < 1e-5


def test_magnetic_b_check_rejects_planted():
    bad = planted(S.magnetic_build_b, "coef = f32(1 / 2 * kk * mu0)", "coef = f32(kk * mu0)")
    assert _check_magnetic_b(bad, (6, 7, 5)) > 1e-3
    bad_pos = planted(S.magnetic_build_b,
                      "pos = tuple((I[a] + (s if a == k else 0)).astype(f32) + f32(0.5)",
                      "pos = tuple((I[a]).astype(f32) + f32(0.5)")
    assert _check_magnetic_b(bad_pos, (6, 7, 5)) > 1e-3
    bad_interp = planted(S.H_interp, "f32(0.0 if a == k else 0.5)", "f32(0.5 if a == k else 0.0)")
    ns = dict(S.magnetic_build_b.__globals__)
    ns["H_interp"] = bad_interp
    src = textwrap.dedent(inspect.getsource(S.magnetic_build_b))
    exec(compile(src, "<planted H_interp>", "exec"), ns)
    assert _check_magnetic_b(ns["magnetic_build_b"], (6, 7, 5)) > 1e-3


def _check_apply_magnetic(apply_fn, res):
    """Linear H_0 = a x, pressure 0 everywhere: a face between a FLUID cell
    and an AIR cell J changes by -scale (p0 + k mu0 |H(J)|^2 / 2) with the
    sign of the side the AIR cell is on, |H(J)| taken at J's centre."""
    dt, rho, dx, p0, a = 0.01, 1000.0, 1.0 / 32, 0.2, 3.0
    ct, vel = _magnetic_scene(res, 61)
    pres = np.zeros(res, f32)
    out = apply_fn(ct, vel, pres, _linear_Hx(res, a), dt, rho, dx, p0, MU0, KK)
    scale = dt / (rho * dx)
    G = _grid(res)
    worst = 0.0
    for k in range(len(res)):
        u = vel[k].astype(np.float64)
        got = out[k].astype(np.float64)
        # face I along k lies between I - e_k and I
        sl = [slice(None)] * len(res)
        sl[k] = slice(1, res[k])
        lo = [slice(None)] * len(res)
        lo[k] = slice(0, res[k] - 1)
        ti, tj = ct[tuple(sl)], ct[tuple(lo)]
        pair = (ti == AIR) & (tj == FLUID)  # fluid below, air above: p_air at I
        Jx = G[0][tuple(sl)]
        pa = p0 + 0.5 * KK * MU0 * (a * (Jx + 0.5)) ** 2
        exp = u[tuple(sl)] - scale * pa
        worst = max(worst, float(np.max(np.abs(got[tuple(sl)] - exp)[pair], initial=0.0)))
        pair2 = (ti == FLUID) & (tj == AIR)  # air below: p_air at I - e_k
        Jx2 = G[0][tuple(lo)]
        pa2 = p0 + 0.5 * KK * MU0 * (a * (Jx2 + 0.5)) ** 2
        exp2 = u[tuple(sl)] + scale * pa2
        worst = max(worst, float(np.max(np.abs(got[tuple(sl)] - exp2)[pair2], initial=0.0)))
    return worst / scale


@pytest.mark.parametrize("res", [(10, 9), (6, 7, 5)])
def test_apply_magnetic_force_closed_form(res):
    assert _check_apply_magnetic(S.apply_magnetic_force, res) < 1e-4


def test_apply_magnetic_check_rejects_planted():
    bad = planted(S.apply_pressure, "v_fa = (u - (scale * (p_air(0) - pj))",
                  "v_fa = (u - (scale * (p_air(-1) - pj))")
    ns = dict(S.apply_magnetic_force.__globals__)
    ns["apply_pressure"] = bad
    src = textwrap.dedent(inspect.getsource(S.apply_magnetic_force))
    exec(compile(src, "<planted apply_pressure>", "exec"), ns)
    assert _check_apply_magnetic(ns["apply_magnetic_force"], (6, 7, 5)) > 1e-3


# ----------------------------------------------------------------------------
# restrict with z != 0 (mgpcg.py:111-118): R (r - A z), scipy assembly
# ----------------------------------------------------------------------------
def _check_restrict_residual(res):
    h = random_hierarchy(res, 2, seed=70, dtype=np.float64)
    gt, Ad, Ax = h["gt"][0], h["Adiag"][0], h["Ax"][0]
    rng = np.random.default_rng(71)
    r = rng.standard_normal(res)
    z = rng.standard_normal(res)
    got = core.restrict(gt, Ad, Ax, r, z)
    A = assemble_matrix(gt, Ad, Ax, fluid_only=False)
    resid = np.where(gt == FLUID, r - (A @ z.ravel()).reshape(res), 0.0)
    exp = np.zeros(got.shape)
    for off in np.ndindex(*([2] * len(res))):
        exp += resid[tuple(slice(o, None, 2) for o in off)]
    return float(np.max(np.abs(got - exp)) / np.max(np.abs(exp)))


@pytest.mark.parametrize("res", [(16, 12), (8, 8, 12)])
def test_restrict_residual_with_nonzero_z(res):
    assert _check_restrict_residual(res) < 1e-12


@pytest.mark.parametrize("old,new", [
    ("T Az = Adiag[idx] * z[idx];", "T Az = (T)0;"),
    ("T res = r[idx] - Az;", "T res = r[idx] + Az;"),
])
def test_restrict_residual_check_rejects_planted(tmp_path_factory, old, new):
    lib = planted_c(tmp_path_factory, [(old, new)])
    with oracle_lib(lib):
        assert _check_restrict_residual((8, 8, 12)) > 1e-3


# ----------------------------------------------------------------------------
# solve (mgpcg.py:150-209) against a textbook PCG with the same V-cycle
# ----------------------------------------------------------------------------
def _pcg_system(res):
    dx = 1.0 / max(res)
    G = _grid(res)
    phi = ((G[1] + 0.5) * dx - 0.6 * res[1] * dx).astype(f32)
    ct = np.where(phi <= 0, FLUID, AIR).astype(np.int8)
    for a in range(len(res)):
        sl = [slice(None)] * len(res)
        sl[a] = 0
        ct[tuple(sl)] = SOLID
        sl[a] = -1
        ct[tuple(sl)] = SOLID
    h = S.hierarchy_from_cell_type(ct, 3, lambda g, l: S.pressure_build_A(g, l, 0.8, False, phi))
    rng = np.random.default_rng(80)
    b = np.where(ct == FLUID, rng.uniform(-1, 1, res), 0.0).astype(f32)
    return h, b


def _textbook_pcg(h64, b, rel_tol=1e-12, abs_tol=1e-14, eps=1e-12, max_iters=500):
    """Preconditioned CG as in the textbooks, x0 = 0, the V-cycle applied
    through the oracle, dots over FLUID cells, stopping when r.r < tol after
    the update (tol on squared norms), iterations = completed non-final bodies."""
    gt = h64["gt"][0]
    fl = (gt == FLUID).ravel()
    A = assemble_matrix(gt, h64["Adiag"][0], h64["Ax"][0])
    r = np.where(fl, b.astype(np.float64).ravel(), 0.0)
    x = np.zeros_like(r)
    dot = lambda u, v: float(np.dot(u[fl], v[fl]))  # noqa: E731
    M = lambda v: core.v_cycle(h64, v.reshape(gt.shape))[0].ravel()  # noqa: E731
    tol = max(abs_tol, rel_tol * dot(r, r))
    z = M(r)
    p = z.copy()
    zr = dot(z, r)
    it = 0
    while it < max_iters:
        Ap = A @ p
        alpha = zr / (dot(p, Ap) + eps)
        x = x + alpha * p
        r = r - alpha * Ap
        if dot(r, r) < tol:
            break
        z = M(r)
        zr_new = dot(z, r)
        beta = zr_new / (zr + eps)
        p = np.where(fl, z + beta * p, 0.0)
        zr = zr_new
        it += 1
    return x.reshape(gt.shape), it


def _check_solve(res):
    h, b = _pcg_system(res)
    h64 = core.to_f64(h)
    xt, it = _textbook_pcg(h64, b)
    x64, o64, _ = core.solve(h64, b.astype(np.float64), precision=core.F64)
    x32, o32, _ = core.solve(h, b, precision=core.PLAIN)
    fl = h["gt"][0] == FLUID
    e64 = rel_l2(x64[fl], xt[fl])
    e32 = rel_l2(x32[fl], xt[fl])
    return max(e64 / 1e-10, e32 / 1e-4, abs(o64["iters"] - it) * 10.0,
               abs(o32["iters"] - it) / 1.0 * 0.5) * 1e-10


@pytest.mark.parametrize("res", [(32, 32), (16, 16, 16)])
def test_solve_equals_textbook_pcg(res):
    assert _check_solve(res) < 1e-10


@pytest.mark.parametrize("old,new", [
    ("double beta = new_zr / (old_zr + eps);", "double beta = new_zr / (pAp + eps);"),
    ("double alpha = old_zr / (pAp + eps);", "double alpha = rr / (pAp + eps);"),
])
def test_solve_check_rejects_planted(tmp_path_factory, old, new):
    lib = planted_c(tmp_path_factory, [(old, new)])
    with oracle_lib(lib):
        assert _check_solve((16, 16, 16)) > 1e-9


# ----------------------------------------------------------------------------
# the core pins of test_oracle_core.py against planted copies of the C oracle
# ----------------------------------------------------------------------------
def _coarsen_ok():
    import itertools

    pats = np.array(list(itertools.product([FLUID, AIR, SOLID], repeat=8)), np.int8)
    fine = np.zeros((2, 2 * len(pats), 2), np.int8)
    for p, pat in enumerate(pats):
        fine[0:2, 2 * p:2 * p + 2, 0:2] = pat.reshape(2, 2, 2)
    rank = {AIR: 2, FLUID: 1, SOLID: 0}
    unrank = {2: AIR, 1: FLUID, 0: SOLID}
    exp = np.array([unrank[max(rank[int(c)] for c in pat)] for pat in pats], np.int8)
    return np.array_equal(core.coarsen(fine).ravel(), exp)


def _compute_ap_ok():
    res = (10, 12, 8)
    hh = random_hierarchy(res, 1, seed=3, dtype=np.float64)
    p = np.random.default_rng(4).standard_normal(res)
    Ap = core.compute_Ap(hh["gt"][0], hh["Adiag"][0], hh["Ax"][0], p)
    A = assemble_matrix(hh["gt"][0], hh["Adiag"][0], hh["Ax"][0])
    fl = hh["gt"][0] == FLUID
    return rel_l2(Ap[fl], (A @ p.ravel()).reshape(res)[fl]) < 1e-12


def _smooth_ok():
    res = (8, 10, 6)
    hh = random_hierarchy(res, 1, seed=5, dtype=np.float64)
    gt, Ad, Ax = hh["gt"][0], hh["Adiag"][0], hh["Ax"][0]
    rng = np.random.default_rng(6)
    r = rng.standard_normal(res)
    z = rng.standard_normal(res)
    core.smooth(gt, Ad, Ax, r, z, 1)
    A = assemble_matrix(gt, Ad, Ax, fluid_only=False)
    resid = r - (A @ z.ravel()).reshape(res)
    G = _grid(res)
    upd = (gt == FLUID) & (((G[0] + G[1] + G[2]) & 1) == 1)
    return float(np.max(np.abs(resid[upd]))) < 1e-12


def _adjoint_ok():
    res = (8, 8, 12)
    hh = random_hierarchy(res, 2, seed=9, dtype=np.float64)
    gt, Ad, Ax = hh["gt"][0], hh["Adiag"][0], hh["Ax"][0]
    rng = np.random.default_rng(10)
    u = rng.standard_normal(res)
    v = rng.standard_normal(tuple(n // 2 for n in res))
    Ru = core.restrict(gt, Ad, Ax, u, np.zeros(res))
    Pv = np.zeros(res)
    core.prolongate(Pv, np.ascontiguousarray(v))
    fl = gt == FLUID
    return abs(float(np.sum(Ru * v)) - float(np.sum(u[fl] * Pv[fl]))) < 1e-10


def _vcycle_symmetric_ok():
    res = (16, 16, 8)
    hh = random_hierarchy(res, 3, seed=13, dtype=np.float64)
    fl = hh["gt"][0] == FLUID
    rng = np.random.default_rng(14)
    a = np.where(fl, rng.standard_normal(res), 0.0)
    b = np.where(fl, rng.standard_normal(res), 0.0)
    Ma, _, _ = core.v_cycle(hh, a)
    Mb, _, _ = core.v_cycle(hh, b)
    ab = float(np.sum(Ma[fl] * b[fl]))
    ba = float(np.sum(a[fl] * Mb[fl]))
    return abs(ab - ba) <= 1e-10 * max(abs(ab), 1.0)


CORE_PINS = {"coarsen": _coarsen_ok, "compute_Ap": _compute_ap_ok, "smooth": _smooth_ok,
             "adjoint": _adjoint_ok, "vcycle_symmetric": _vcycle_symmetric_ok}

POST_SMOOTH = """      orc_prolongate_##SFX(dim, n, z[l], z[l + 1]);                                             \\
      for (int i = 0; i < nu; ++i) {                                                            \\
        orc_smooth_##SFX(dim, n, gt[l], Adiag[l], Ax[l], r[l], z[l], 1);                        \\
        orc_smooth_##SFX(dim, n, gt[l], Adiag[l], Ax[l], r[l], z[l], 0);                        \\"""

CORE_PLANTS = [
    ("coarsen", "int8_t t = tot_air > 0 ? AIR : (tot_fluid > 0 ? FLUID : SOLID);",
     "int8_t t = tot_fluid > 0 ? FLUID : (tot_air > 0 ? AIR : SOLID);"),
    ("compute_Ap", "T p = Ax[(idx - s) * g->dim + a] * x[idx - s];",
     "T p = Ax[idx * g->dim + a] * x[idx - s];"),
    ("smooth", "if (((c0 + c1 + c2) & 1) != phase) continue;",
     "if (((c0 + c1) & 1) != phase) continue;"),
    ("adjoint", "z[idx] = z[idx] + zc[cidx];", "z[idx] = z[idx] + zc[cidx] + zc[cidx];"),
    ("vcycle_symmetric", POST_SMOOTH,
     POST_SMOOTH.replace("z[l], 1);", "z[l], 2);").replace("z[l], 0);", "z[l], 1);")
     .replace("z[l], 2);", "z[l], 0);")),
]


@pytest.mark.parametrize("pin", sorted(CORE_PINS))
def test_core_pin_holds_on_the_oracle(pin):
    assert CORE_PINS[pin]()


@pytest.mark.parametrize("pin,old,new", CORE_PLANTS, ids=[p[0] for p in CORE_PLANTS])
def test_core_pin_rejects_planted_c_oracle(tmp_path_factory, pin, old, new):
    lib = planted_c(tmp_path_factory, [(old, new)])
    with oracle_lib(lib):
        assert not CORE_PINS[pin]()


# ----------------------------------------------------------------------------
# PressureProjectStrategy.build_A (pressure_project.py:45-64), GFM off
# ----------------------------------------------------------------------------
def _check_pressure_dirichlet(build_A, res=(10, 8, 12), modes=(2, 1, 3)):
    """All-FLUID interior in an AIR ring, GFM off: A = sA times the Dirichlet
    Laplacian, so A v = lambda v for v = prod sin(pi k_a i_a / (m_a + 1))."""
    sA = 0.375
    ct = np.full(res, AIR, np.int8)
    ct[tuple(slice(1, n - 1) for n in res)] = FLUID
    Ad, Ax = build_A(ct, 0, sA, ghost_fluid_method=False)
    Ad, Ax = Ad.astype(np.float64), Ax.astype(np.float64)
    G = _grid(res)
    v = np.ones(res)
    lam = 0.0
    for a, (n, k) in enumerate(zip(res, modes)):
        m = n - 2
        v = v * np.sin(np.pi * k * G[a] / (m + 1))
        lam += 2 - 2 * np.cos(np.pi * k / (m + 1))
    lam *= sA
    v = np.where(ct == FLUID, v, 0.0)
    A = assemble_matrix(ct, Ad, Ax)
    Av = (A @ v.ravel()).reshape(res)
    fl = ct == FLUID
    return rel_l2(Av[fl], lam * v[fl])


def test_pressure_A_dirichlet_eigenpair():
    assert _check_pressure_dirichlet(S.pressure_build_A) < 1e-12


@pytest.mark.parametrize("old,new", [
    ("Adiag = np.where(m_a, (Adiag + sA).astype(f32), Adiag)", "Adiag = Adiag"),
    ("Ax[..., k] = np.where(m_f, -sA, Ax[..., k])", "Ax[..., k] = np.where(m_f, sA, Ax[..., k])"),
    ("if s > 0:\n                Ax[..., k]", "if s < 0:\n                Ax[..., k]"),
])
def test_pressure_A_check_rejects_planted(old, new):
    bad = planted(S.pressure_build_A, old, new)
    assert _check_pressure_dirichlet(bad) > 1e-3
