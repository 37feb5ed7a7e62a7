"""Seeded synthetic Poisson systems: the workloads the solver is measured on.

Every generator is host-side numpy (``default_rng(seed)``, PCG64) so the same
seed gives the same system on any machine.  Cell index i maps to the world
coordinate (i + 0.5) * dx with one dx for all axes; axis 1 is "y".

A system is a dict with ``kind`` in {"pressure", "volume", "psi", "surface"},
``dim``, ``res``, solver parameters (``L``, ``nu``, ``bottom``) and the fields
its strategy reads.  ``make_solver_inputs`` moves it to the device.

Workloads (sizes and seeds):
  c1_2d(seed=0)            2D 128^2 pressure, SOLID border, FLUID below y=0.35
                           or inside the disc (0.5, 0.7) r=0.12, no GFM, p0=0,
                           L=4, nu=2, bottom=10.
  c2_3d(seed=1)            3D 128^3 pressure with ghost-fluid coefficients:
                           phi = min(y - 0.3, |x - (0.5,0.65,0.5)| - 0.15),
                           GFM clamp theta >= 0.01 (c_max = 100), L=5.
  c3_256(seed=2)           3D 256^3 pool (y < 0.3) plus 8 spheres, centres
                           U[0.15,0.85]^3, radii U[0.05,0.12], L=6.
  c4_batch(seeds 100..163) 64 systems, sides drawn per axis from
                           {32,48,64,96,128}, kinds cycling psi / pressure /
                           surface tension / volume control, L=5.
  c5_request(seed)         side lengths log-uniform in [64,256] per axis,
                           rounded to a multiple of 16, pool+sphere pressure.
"""
import math

import numpy as np

FLUID, AIR, SOLID = 0, 1, 2
RHO = 1000.0  # fluid_simulator.py:32
DT = 1.25e-2  # fluid_simulator.py:29


def centres(res, dx):
    return np.meshgrid(*[(np.arange(n, dtype=np.float64) + 0.5) * dx for n in res], indexing="ij")


def pool_spheres_phi(res, dx, pool_y, spheres):
    """Signed distance to the union of the half-space y < pool_y and spheres."""
    X = centres(res, dx)
    phi = X[1] - pool_y
    for c, r in spheres:
        d = np.sqrt(sum((X[a] - c[a]) ** 2 for a in range(len(res)))) - r
        phi = np.minimum(phi, d)
    return phi.astype(np.float32)


def cell_types(phi, solid_border=True):
    """FLUID where phi <= 0 (fluid_simulator.py:151-159), AIR elsewhere, SOLID
    one-cell border (fluid_simulator.py:337-341); int32 as the simulator holds it."""
    ct = np.where(phi <= 0, FLUID, AIR).astype(np.int32)
    if solid_border:
        for a in range(phi.ndim):
            sl = [slice(None)] * phi.ndim
            sl[a] = 0
            ct[tuple(sl)] = SOLID
            sl[a] = -1
            ct[tuple(sl)] = SOLID
    return ct


def mac_velocity(res, rng):
    """Seeded U[-1, 1] MAC components, component k of shape res + e_k."""
    out = []
    for k in range(len(res)):
        shape = [n + (1 if a == k else 0) for a, n in enumerate(res)]
        out.append(rng.uniform(-1.0, 1.0, shape).astype(np.float32))
    return out


def _pressure(res, dx, phi, ct, vel, L, gfm, c_max, p0, seed, name):
    sA = DT / (RHO * dx * dx)
    return {
        "name": name, "kind": "pressure", "seed": seed, "dim": len(res), "res": list(res),
        "L": L, "nu": 2, "bottom": 10, "dx": dx, "dt": DT, "rho": RHO, "scale_A": sA,
        "scale_b": 1.0 / dx, "p0": p0, "gfm": gfm, "c_max": c_max, "phi": phi,
        "cell_type": ct, "velocity": vel,
    }


def c1_2d(seed=0, n=128):
    res = [n, n]
    dx = 1.0 / n
    rng = np.random.default_rng(seed)
    X = centres(res, dx)
    disc = np.sqrt((X[0] - 0.5) ** 2 + (X[1] - 0.7) ** 2) - 0.12
    phi = np.minimum(X[1] - 0.35, disc).astype(np.float32)
    ct = cell_types(phi)
    vel = mac_velocity(res, rng)
    return _pressure(res, dx, phi, ct, vel, 4, False, 1e3, 0.0, seed, f"c1_2d_{n}")


def c2_3d(seed=1, n=128, L=5):
    res = [n, n, n]
    dx = 1.0 / n
    rng = np.random.default_rng(seed)
    phi = pool_spheres_phi(res, dx, 0.3, [((0.5, 0.65, 0.5), 0.15)])
    ct = cell_types(phi)
    vel = mac_velocity(res, rng)
    return _pressure(res, dx, phi, ct, vel, L, True, 100.0, 1e-3, seed, f"c2_3d_{n}")


def c3_256(seed=2, n=256, L=6):
    res = [n, n, n]
    dx = 1.0 / n
    rng = np.random.default_rng(seed)
    cs = rng.uniform(0.15, 0.85, (8, 3))
    rs = rng.uniform(0.05, 0.12, 8)
    phi = pool_spheres_phi(res, dx, 0.3, [(tuple(c), r) for c, r in zip(cs, rs)])
    ct = cell_types(phi)
    vel = mac_velocity(res, rng)
    return _pressure(res, dx, phi, ct, vel, L, True, 100.0, 1e-3, seed, f"c3_{n}")


def pool_sphere_system(res, seed, L=5, kind="pressure"):
    """Pool + one seeded sphere on an arbitrary box (C4/C5 scenes)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / max(res)
    ext = [n * dx for n in res]
    c = tuple(rng.uniform(0.3, 0.7) * e for e in ext)
    r = rng.uniform(0.1, 0.2) * min(ext)
    phi = pool_spheres_phi(res, dx, 0.3 * ext[1], [(c, r)])
    ct = cell_types(phi)
    vel = mac_velocity(res, rng)
    sysd = _pressure(res, dx, phi, ct, vel, L, True, 100.0, 1e-3, seed,
                     f"{kind}_{'x'.join(map(str, res))}_s{seed}")
    if kind == "pressure":
        return sysd
    if kind == "volume":
        # the second volume-controlled solve of a run: a seeded divergence constant
        sysd["kind"] = "volume"
        sysd["c"] = float(rng.uniform(-0.5, 0.5))
        return sysd
    if kind == "psi":
        chi = np.where(phi <= 0, 2.0, 0.0).astype(np.float32)  # chi = 2 in fluid, 0 in air
        H_ext = []
        for k in range(len(res)):
            shape = [n + (1 if a == k else 0) for a, n in enumerate(res)]
            H_ext.append((1.2e4 * rng.uniform(-1.0, 1.0, shape)).astype(np.float32))
        sysd.update(kind="psi", chi=chi, H_ext=H_ext)
        return sysd
    if kind == "surface":
        sysd.update(kind="surface", d=int(rng.integers(0, len(res))), sigma=7.28e-2)
        return sysd
    raise ValueError(kind)


C4_SIDES = (32, 48, 64, 96, 128)
C4_KINDS = ("psi", "pressure", "surface", "volume")


def c4_batch(first_seed=100, count=64, L=5):
    out = []
    for i in range(count):
        seed = first_seed + i
        rng = np.random.default_rng(seed)
        res = [int(rng.choice(C4_SIDES)) for _ in range(3)]
        out.append(pool_sphere_system(res, seed, L=L, kind=C4_KINDS[i % 4]))
    return out


def c5_request(seed, lo=64, hi=256, L=5):
    rng = np.random.default_rng(seed)
    res = []
    for _ in range(3):
        side = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        res.append(int(max(lo, min(hi, round(side / 16) * 16))))
    return pool_sphere_system(res, seed, L=L, kind="pressure")
