import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "taichi-ferrofluid_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


EMU_LIB = os.path.join(ROOT, "tests", "cpu_emu", "libmgpcg_emu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmgpcg.so)")


@pytest.fixture(scope="session", params=[pytest.param("cuda:0", marks=pytest.mark.gpu), "emu"])
def dev(request):
    """Device of a parity test: the MI355X through libmgpcg.so ("cuda:0", marked
    gpu), or the CPU emulation build of the same HIP sources ("emu" -> torch
    CPU tensors), which runs every kernel and the engine bounds-checked on the
    host.  One process loads one of the two libraries."""
    from mgpcg_amd import _lib

    if request.param == "emu":
        if _lib._lib is not None and _lib._path != EMU_LIB:
            pytest.skip("libmgpcg.so already loaded in this process")
        sys.path.insert(0, os.path.join(ROOT, "tests", "cpu_emu"))
        import build_emu

        build_emu.build_if_stale()
        _lib.use_library(EMU_LIB)
        return "cpu"
    if _lib._lib is not None and _lib._path == EMU_LIB:
        pytest.skip("the CPU emulation build is loaded in this process")
    return "cuda:0"


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
