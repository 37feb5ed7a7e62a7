"""The C-ABI library loads and exports exactly what include/mgpcg.h declares.

CPU-only: the library is opened with dlopen and inspected; no function that
touches a device is called.
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mgpcg.h")
LIB = os.path.join(ROOT, "taichi-ferrofluid_amd", "libmgpcg.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mgpcg_[A-Za-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def exported_functions():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True,
                         text=True).stdout
    names = set()
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[1] == "T" and parts[2].startswith("mgpcg_"):
            names.add(parts[2])
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is missing: run __graft_entry__.build()")
    return ctypes.CDLL(LIB)


def test_header_declares_the_binding_table():
    from mgpcg_amd import _lib

    assert declared_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_nothing_undeclared():
    assert exported_functions() == declared_functions()


def test_version_matches_header(lib):
    m = re.search(r"#define MGPCG_ABI_VERSION (\d+)", open(HEADER).read())
    lib.mgpcg_version.restype = ctypes.c_int
    assert lib.mgpcg_version() == int(m.group(1))


def test_constants_match_header():
    """The Python binding's constants are the header's #defines."""
    from mgpcg_amd import _lib

    defs = dict(re.findall(r"#define (MGPCG_[A-Z0-9_]+)\s+(-?\d+)", open(HEADER).read()))
    pairs = {"PLAIN": "MGPCG_PLAIN", "MIXED": "MGPCG_MIXED", "F64": "MGPCG_F64",
             "RUNNING": "MGPCG_RUNNING", "CONVERGED": "MGPCG_CONVERGED",
             "MAX_ITERS": "MGPCG_MAX_ITERS", "BREAKDOWN": "MGPCG_BREAKDOWN",
             "FIELD_R": "MGPCG_FIELD_R", "FIELD_Z": "MGPCG_FIELD_Z", "FIELD_P": "MGPCG_FIELD_P",
             "FIELD_AP": "MGPCG_FIELD_AP", "FIELD_X": "MGPCG_FIELD_X",
             "FIELD_P_NEXT": "MGPCG_FIELD_P_NEXT", "OP_SMOOTH": "MGPCG_OP_SMOOTH",
             "OP_SMOOTH_PAIR_FROM_ZERO": "MGPCG_OP_SMOOTH_PAIR_FROM_ZERO",
             "OP_RESTRICT": "MGPCG_OP_RESTRICT", "OP_PROLONGATE": "MGPCG_OP_PROLONGATE",
             "OP_VCYCLE": "MGPCG_OP_VCYCLE", "OP_COMPUTE_AP": "MGPCG_OP_COMPUTE_AP",
             "OP_PROLONG_SMOOTH1": "MGPCG_OP_PROLONG_SMOOTH1",
             "OPT_COARSE_MAX_SLOTS": "MGPCG_OPT_COARSE_MAX_SLOTS",
             "OPT_MAX_GRID": "MGPCG_OPT_MAX_GRID", "OPT_FUSE_FIRST": "MGPCG_OPT_FUSE_FIRST",
             "OPT_FUSE_PROLONG": "MGPCG_OPT_FUSE_PROLONG", "OPT_FUSE_P": "MGPCG_OPT_FUSE_P",
             "OPT_COMPACT": "MGPCG_OPT_COMPACT", "OPT_COARSE_LDS": "MGPCG_OPT_COARSE_LDS",
             "OPT_RESTRICT_ROWS": "MGPCG_OPT_RESTRICT_ROWS",
             "OPT_PROBE_EVERY": "MGPCG_OPT_PROBE_EVERY",
             "OPT_SERVER_DEPTH": "MGPCG_OPT_SERVER_DEPTH",
             "OPT_SERVER_STREAMS": "MGPCG_OPT_SERVER_STREAMS", "OPT_GRAPH": "MGPCG_OPT_GRAPH",
             "OPT_RESTRICT_COARSE": "MGPCG_OPT_RESTRICT_COARSE"}
    for py, c in pairs.items():
        assert getattr(_lib, py) == int(defs[c]), (py, c)
    # every #define of the header is mirrored
    assert set(defs) - set(pairs.values()) <= {"MGPCG_ABI_VERSION"}


def test_result_struct_layout():
    from mgpcg_amd import _lib

    assert ctypes.sizeof(_lib.MgpcgResult) == 8 + 6 * 8


def test_links_the_hip_runtime_by_soname():
    """The library resolves libamdhip64 by soname, so the process's HIP
    runtime (the one torch mapped) is the one it uses."""
    out = subprocess.run(["readelf", "-d", LIB], check=True, capture_output=True,
                         text=True).stdout
    needed = re.findall(r"\(NEEDED\).*\[(.+)\]", out)
    hip = [n for n in needed if "amdhip64" in n]
    assert hip == ["libamdhip64.so.7"], needed
