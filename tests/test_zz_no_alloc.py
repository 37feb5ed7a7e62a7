"""The allocation-free entry point mgpcg_calc_volume_static against
mgpcg_calc_volume (the same kernels with a caller-provided partials buffer),
through the C-ABI: bit-identical volumes on two calls in a row, the second
on a smaller grid (the shared scratch array keeps the first call's partials
beyond the second's), and a grid beyond the array's limit rejected before
any launch."""
import ctypes

import numpy as np
import pytest
import torch

from mgpcg_amd import scenes


def _volume(lib, dev, res, phi, static):
    from mgpcg_amd import _lib

    out = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    stream = _lib.stream_handle(dev)
    if static:
        _lib.check(lib.mgpcg_calc_volume_static(stream, len(res), _lib.int3(res), _lib.ptr(phi),
                                                1.0 / max(res), _lib.ptr(out)))
    else:
        nb = (int(np.prod(res)) + 255) // 256
        partials = torch.zeros(nb, dtype=torch.float64, device=dev)
        _lib.check(lib.mgpcg_calc_volume(stream, len(res), _lib.int3(res), _lib.ptr(phi),
                                         1.0 / max(res), _lib.ptr(partials), nb, _lib.ptr(out)))
    _lib.synchronize(dev)
    return float(out.item())


@pytest.mark.gpu
def test_calc_volume_static_equals_calc_volume():
    from mgpcg_amd import _lib

    dev = "cuda:0"
    lib = _lib.load()
    for res in ([96, 128, 80], [32, 48, 16], [64, 64]):  # the second is smaller
        dx = 1.0 / max(res)
        if len(res) == 3:
            phi = scenes.pool_spheres_phi(res, dx, 0.3, [((0.5, 0.6, 0.5), 0.2)])
        else:
            X = scenes.centres(res, dx)
            phi = np.minimum(X[1] - 0.35, np.sqrt((X[0] - 0.5) ** 2 + (X[1] - 0.7) ** 2) - 0.12)
        phi_d = torch.from_numpy(np.ascontiguousarray(phi, np.float32)).to(dev)
        ref = _volume(lib, dev, res, phi_d, static=False)
        got = _volume(lib, dev, res, phi_d, static=True)
        print(f"case calc_volume_static {res}: {got!r} vs {ref!r}", flush=True)
        assert got == ref and ref > 0.0


@pytest.mark.gpu
def test_calc_volume_static_rejects_grids_beyond_its_limit():
    from mgpcg_amd import _lib

    lib = _lib.load()
    out = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    phi = torch.zeros(1, dtype=torch.float32, device="cuda:0")  # never read: rejected first
    rc = lib.mgpcg_calc_volume_static(_lib.stream_handle("cuda:0"), 3, _lib.int3([512, 512, 512]),
                                      _lib.ptr(phi), 1.0 / 512, _lib.ptr(out))
    assert rc == -1
    assert b"allocation-free limit" in lib.mgpcg_last_error()
