"""ctypes binding of libmgpcg.so (the HIP library; C-ABI in include/mgpcg.h).

torch is imported first so that the library's libamdhip64.so.7 dependency
resolves to the HIP runtime torch already mapped (one runtime per process).
There is no fallback: a missing or unloadable library raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede loading the HIP library)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libmgpcg.so")  # taichi-ferrofluid_amd/

PLAIN, MIXED, F64 = 0, 1, 2
RUNNING, CONVERGED, MAX_ITERS, BREAKDOWN = 0, 1, 2, 3
FIELD_R, FIELD_Z, FIELD_P, FIELD_AP, FIELD_X, FIELD_P_NEXT = 0, 1, 2, 3, 4, 5
(OP_SMOOTH, OP_SMOOTH_PAIR_FROM_ZERO, OP_RESTRICT, OP_PROLONGATE, OP_VCYCLE, OP_COMPUTE_AP,
 OP_PROLONG_SMOOTH1) = range(7)
(OPT_COARSE_MAX_SLOTS, OPT_MAX_GRID, OPT_FUSE_FIRST, OPT_FUSE_PROLONG, OPT_FUSE_P,
 OPT_COMPACT, OPT_COARSE_LDS, OPT_RESTRICT_ROWS, OPT_PROBE_EVERY,
 OPT_SERVER_DEPTH, OPT_SERVER_STREAMS, OPT_GRAPH,
 OPT_RESTRICT_COARSE) = 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


class MgpcgResult(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("r0r0", ctypes.c_double),
        ("rr", ctypes.c_double),
        ("true_rr", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes); every symbol include/mgpcg.h declares
SIGNATURES = {
    "mgpcg_create": (_I, [_I, _P, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "mgpcg_destroy": (_I, [_P]),
    "mgpcg_set_stream": (_I, [_P, _P]),
    "mgpcg_set_option": (_I, [_P, _I, _I]),
    "mgpcg_get_option": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "mgpcg_info": (_I, [_P, _I, _P, _P]),
    "mgpcg_assemble_pressure": (_I, [_P, _P, _P, _P, _D, _D, _D, _I, _D, _I, _D]),
    "mgpcg_pool_bytes": (_I, [_I, _P, _I, _I, _P]),
    "mgpcg_load_level": (_I, [_P, _I, _P, _P, _P]),
    "mgpcg_load_rhs": (_I, [_P, _P]),
    "mgpcg_level_coupling": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_float)]),
    "mgpcg_level_flags": (_I, [_P, _I, _P, _P]),
    "mgpcg_level_order": (_I, [_P, _I, _P, _P]),
    "mgpcg_coarsen": (_I, [_P, _I, _P, _P, _P]),
    "mgpcg_solve": (_I, [_P, _I, _D, _D, _D, _I, ctypes.POINTER(MgpcgResult)]),
    "mgpcg_solve_hostsync": (_I, [_P, _I, _D, _D, _D, _I, ctypes.POINTER(MgpcgResult)]),
    "mgpcg_solve_reference_structured": (_I, [_P, _I, _D, _D, _D,
                                              ctypes.POINTER(MgpcgResult)]),
    "mgpcg_store_x": (_I, [_P, _P, _P]),
    "mgpcg_debug_set": (_I, [_P, _I, _I, _P]),
    "mgpcg_debug_get": (_I, [_P, _I, _I, _P]),
    "mgpcg_debug_op": (_I, [_P, _I, _I, _I]),
    "mgpcg_time_op": (_I, [_P, _I, _I, _I, _I, ctypes.POINTER(_D)]),
    "mgpcg_pressure_build_b": (_I, [_P, _I, _P, _P, _P, _P, _D, _D, _D, _I, _D, _P]),
    "mgpcg_pressure_build_A": (_I, [_P, _I, _P, _I, _P, _P, _D, _I, _D, _P, _P]),
    "mgpcg_magnetic_build_b": (_I, [_P, _I, _P, _P, _P, _P, _D, _D, _D, _D, _D, _P]),
    "mgpcg_calc_volume": (_I, [_P, _I, _P, _P, _D, _P, _I, _P]),
    "mgpcg_calc_volume_static": (_I, [_P, _I, _P, _P, _D, _P]),
    "mgpcg_add_c": (_I, [_P, _I, _P, _P, _D, _P]),
    "mgpcg_potential_build_b": (_I, [_P, _I, _P, _P, _P, _P]),
    "mgpcg_potential_build_A": (_I, [_P, _I, _P, _I, _D, _P, _P, _P]),
    "mgpcg_calc_n": (_I, [_P, _I, _P, _I, _P, _D, _P]),
    "mgpcg_calc_kappa": (_I, [_P, _I, _P, _P, _D, _P]),
    "mgpcg_surface_build_b": (_I, [_P, _I, _P, _I, _P, _P, _P, _P, _D, _D, _D, _P]),
    "mgpcg_surface_build_A": (_I, [_P, _I, _P, _I, _I, _P, _D, _D, _D, _P, _P]),
    "mgpcg_apply_pressure": (_I, [_P, _I, _P, _P, _P, _P, _P, _D, _D, _D, _D, _I, _D]),
    "mgpcg_apply_magnetic_force": (_I, [_P, _I, _P, _P, _P, _P, _P, _D, _D, _D, _D, _D, _D]),
    "mgpcg_apply_markers": (_I, [_P, _I, _P, _P, _P]),
    "mgpcg_init_boundary": (_I, [_P, _I, _P, _P]),
    "mgpcg_redistance": (_I, [_P, _I, _P, _D, _I, _P, _P, _P, _P]),
    "mgpcg_extrap_velocity": (_I, [_P, _I, _P, _P, _P, _I, _P, _P, _P]),
    "mgpcg_batch_create": (_I, [ctypes.POINTER(_P)]),
    "mgpcg_batch_destroy": (_I, [_P]),
    "mgpcg_batch_set_stream": (_I, [_P, _P]),
    "mgpcg_batch_solve": (_I, [_P, _P, _I, _I, _D, _D, _D, _I, _P]),
    "mgpcg_server_create": (_I, [_I, _D, _D, _D, _I, ctypes.POINTER(_P)]),
    "mgpcg_server_destroy": (_I, [_P]),
    "mgpcg_server_set_stream": (_I, [_P, _P]),
    "mgpcg_server_submit": (_I, [_P, _P, ctypes.c_int64]),
    "mgpcg_server_step": (_I, [_P, ctypes.POINTER(_I)]),
    "mgpcg_server_poll": (_I, [_P, _P, _P, _I, ctypes.POINTER(_I)]),
    "mgpcg_server_pending": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                  ctypes.POINTER(_I)]),
    "mgpcg_server_probe": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_D),
                                ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "mgpcg_batch_probe": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_D),
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "mgpcg_graph_counts": (_I, [_P, _P, _P, ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_int64)]),
    "mgpcg_last_error": (ctypes.c_char_p, []),
    "mgpcg_version": (_I, []),
}

_lib = None


class MgpcgError(RuntimeError):
    pass


_path = LIB_PATH


def use_library(path):
    """Select another build of the same C-ABI before the first load (the
    tests' CPU emulation build); the product always uses libmgpcg.so."""
    global _path
    if _lib is not None and path != _path:
        raise MgpcgError("the library is already loaded")
    _path = path


def load():
    """Load libmgpcg.so (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_path):
            raise MgpcgError(
                f"{_path} is missing; build it with `python taichi-ferrofluid_amd/build.py` "
                "(or __graft_entry__.build())")
        lib = ctypes.CDLL(_path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != 0:
        raise MgpcgError(load().mgpcg_last_error().decode(errors="replace"))
    return rc


def ptr(t):
    """Device (or host) address of a contiguous torch tensor."""
    if t is None:
        return None
    assert t.is_contiguous(), "tensors crossing the C-ABI must be contiguous"
    return ctypes.c_void_p(t.data_ptr())


def ptr_array(tensors):
    """A C array of the tensors' addresses (None -> NULL)."""
    return (ctypes.c_void_p * len(tensors))(*[None if t is None else t.data_ptr()
                                              for t in tensors])


def int3(vals):
    v = list(vals) + [0] * (3 - len(vals))
    return (ctypes.c_int32 * 3)(*v)


_dev_index = {}


def stream_handle(device=None):
    """torch's current HIP stream of `device` (NULL for CPU tensors, which only
    the CPU emulation build accepts)."""
    idx = _dev_index.get(device)
    if idx is None:
        d = torch.device(device) if device is not None else None
        if d is not None and d.type == "cpu":
            idx = -1
        else:
            idx = d.index if d is not None and d.index is not None else torch.cuda.current_device()
        if device is not None:  # the current device may change between calls
            _dev_index[device] = idx
    if idx < 0:
        return ctypes.c_void_p(None)
    # the raw stream of torch's current stream on that device (torch.cuda.current_stream
    # without building a Stream object)
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(idx))


def synchronize(device):
    if torch.device(device).type != "cpu":
        torch.cuda.synchronize(device)


def graph_counts(solver=None, batch=None, server=None):
    """(captures, graph launches) of exactly one handle's engines."""
    c, n = ctypes.c_int64(), ctypes.c_int64()
    check(load().mgpcg_graph_counts(solver, batch, server, ctypes.byref(c), ctypes.byref(n)))
    return c.value, n.value


def hip_runtimes_mapped():
    """Paths of every libamdhip64 mapped into this process (should be one)."""
    paths = set()
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                paths.add(line.split()[-1])
    return sorted(paths)
