"""Operator assembly in the HIP library vs the numpy strategies.

Pressure / magnetic / volume-control arithmetic has no transcendental
functions and must match bit for bit; psi (sample only) too.  Surface tension
and calc_volume go through sin/cos (level_set.py:24-37), whose FP32 results
may differ by an ulp between math libraries: those compare within 1e-5
relative.  Runs on the MI355X (`dev` = "cuda:0", marked gpu) and on the CPU
emulation build (`dev` = "cpu").
"""
import numpy as np
import pytest
import torch

from mgpcg_amd import scenes
from oracle import systems as osys
from oracle import strategies as S


def _assemble(sysd, dev):
    from mgpcg_amd import _lib
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    d = systems.upload(sysd, dev)
    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], 2, 10, device=dev)
    systems.assemble(s, d)
    _lib.synchronize(dev)
    return s


def _cmp(s, h, b, exact):
    L = len(h["gt"])
    for l in range(L):
        assert np.array_equal(s.grid_type[l].cpu().numpy(), h["gt"][l]), f"gt level {l}"
        for name, got, ref in (("Adiag", s.Adiag[l], h["Adiag"][l]), ("Ax", s.Ax[l], h["Ax"][l])):
            g = got.cpu().numpy()
            if exact:
                assert np.array_equal(g, ref), f"{name} level {l}"
            else:
                scale = max(np.abs(ref).max(), 1e-30)
                assert np.abs(g - ref).max() <= 1e-5 * scale, f"{name} level {l}"
    g = s.b.cpu().numpy()
    if exact:
        assert np.array_equal(g, b)
    else:
        assert np.abs(g - b).max() <= 1e-5 * max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("make", [
    lambda: scenes.c1_2d(seed=0, n=64),
    lambda: scenes.c2_3d(seed=1, n=32, L=3),
    lambda: scenes.pool_sphere_system([48, 32, 64], 5, L=4, kind="pressure"),
    lambda: scenes.pool_sphere_system([32, 48, 32], 6, L=4, kind="volume"),
    lambda: scenes.pool_sphere_system([32, 32, 48], 7, L=4, kind="psi"),
], ids=["c1_2d_64", "c2_32", "pressure_48x32x64", "volume_32x48x32", "psi_32x32x48"])
def test_assembly_bit_exact(dev, make):
    sysd = make()
    s = _assemble(sysd, dev)
    h, b = osys.assemble(sysd)
    _cmp(s, h, b, exact=True)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_surface_tension_assembly(dev, d):
    sysd = scenes.pool_sphere_system([32, 32, 32], 8, L=3, kind="surface")
    sysd["d"] = d
    s = _assemble(sysd, dev)
    h, b = osys.assemble(sysd)
    _cmp(s, h, b, exact=False)


def test_magnetic_b_bit_exact(dev):
    from mgpcg_amd import magnetic_force
    from mgpcg_amd import mgpcg
    from mgpcg_amd import systems

    sysd = scenes.c2_3d(seed=3, n=32, L=3)
    rng = np.random.default_rng(4)
    H = []
    for k in range(3):
        shape = [n + (1 if a == k else 0) for a, n in enumerate(sysd["res"])]
        H.append((1e3 * rng.uniform(-1, 1, shape)).astype(np.float32))
    d = systems.upload(sysd, dev)
    Hd = [torch.from_numpy(h).to(dev) for h in H]
    s = mgpcg.MGPCGPoissonSolver(3, sysd["res"], 3, 2, 10, device=dev)
    mu0, kk = 4 * np.pi * 1e-7, 5
    st = magnetic_force.MagneticForceStrategy(3, d["velocity"], True, d["phi"], 1e-3, mu0, kk, Hd)
    st.scale_A, st.scale_b = sysd["scale_A"], sysd["scale_b"]
    s.reinitialize(d["cell_type"], st)
    ref = S.magnetic_build_b(np.ascontiguousarray(sysd["cell_type"], np.int8), sysd["velocity"], H,
                             sysd["scale_A"], sysd["scale_b"], 1e-3, mu0, kk)
    assert np.array_equal(s.b.cpu().numpy(), ref)


def test_calc_volume(dev):
    from mgpcg_amd import volume_control
    from mgpcg_amd.level_set import LevelSet

    sysd = scenes.c2_3d(seed=1, n=32, L=3)
    phi = torch.from_numpy(sysd["phi"]).to(dev)
    ls = LevelSet(3, sysd["res"], sysd["dx"], phi=phi, device=dev)
    st = volume_control.PressureProjectWithVolumeControlStrategy(
        3, [], True, phi, 0.0, ls, sysd["dt"])
    vol = st.calc_volume()
    ref = S.calc_volume(sysd["phi"], sysd["dx"])
    assert abs(vol - ref) <= 1e-5 * abs(ref)
