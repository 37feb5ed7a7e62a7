"""Operator assembly on the device vs the numpy strategies (GPU).

Pressure / magnetic / volume-control arithmetic has no transcendental
functions and must match bit for bit.  psi is bit-exact too (sample only).
Surface tension and calc_volume go through sin/cos (level_set.py:24-37),
whose FP32 results may differ by an ulp between the GPU math library and
numpy: those compare within 1e-5 relative.
"""
import numpy as np
import pytest
import torch

import scenes
from oracle import core
from oracle import systems as osys
from oracle import strategies as S

pytestmark = pytest.mark.gpu


def _assemble_gpu(sysd):
    import mgpcg
    import systems

    dev = systems.upload(sysd, "cuda:0")
    s = mgpcg.MGPCGPoissonSolver(sysd["dim"], sysd["res"], sysd["L"], 2, 10, device="cuda:0")
    systems.assemble(s, dev)
    torch.cuda.synchronize()
    return s, dev


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
def test_assembly_bit_exact(make):
    sysd = make()
    s, _ = _assemble_gpu(sysd)
    h, b = osys.assemble(sysd)
    _cmp(s, h, b, exact=True)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_surface_tension_assembly(d):
    sysd = scenes.pool_sphere_system([32, 32, 32], 8, L=3, kind="surface")
    sysd["d"] = d
    s, _ = _assemble_gpu(sysd)
    h, b = osys.assemble(sysd)
    _cmp(s, h, b, exact=False)


def test_calc_volume():
    import volume_control
    from level_set import LevelSet

    sysd = scenes.c2_3d(seed=1, n=32, L=3)
    phi = torch.from_numpy(sysd["phi"]).cuda()
    ls = LevelSet(3, sysd["res"], sysd["dx"], phi=phi)
    st = volume_control.PressureProjectWithVolumeControlStrategy(
        3, [], True, phi, 0.0, ls, sysd["dt"])
    vol = st.calc_volume()
    ref = S.calc_volume(sysd["phi"], sysd["dx"])
    assert abs(vol - ref) <= 1e-5 * abs(ref)
