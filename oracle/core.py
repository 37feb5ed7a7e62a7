"""ctypes bindings of the C oracle (oracle/mgpcg_oracle.c).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Arrays are numpy, canonical C-order per level (shape ``res >> l``); ``Ax`` is
``[*res_l, dim]`` (mgpcg.py:34-35).  Cell types are int8 FLUID=0, AIR=1,
SOLID=2 (utils.py:4-6, mgpcg.py:29).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PLAIN, MIXED, F64 = 0, 1, 2
FLUID, AIR, SOLID = 0, 1, 2

_lib = None


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int),
        ("status", ctypes.c_int),
        ("r0r0", ctypes.c_double),
        ("rr", ctypes.c_double),
        ("true_rr", ctypes.c_double),
        ("tol", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build

            build.build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_long


def _declare(L):
    L.orc_coarsen.argtypes = [_I, _P, _P, _P]
    for sfx in ("f32", "f64"):
        getattr(L, f"orc_smooth_{sfx}").argtypes = [_I, _P, _P, _P, _P, _P, _P, _I]
        getattr(L, f"orc_restrict_{sfx}").argtypes = [_I, _P, _P, _P, _P, _P, _P, _P]
        getattr(L, f"orc_prolongate_{sfx}").argtypes = [_I, _P, _P, _P]
        getattr(L, f"orc_compute_Ap_{sfx}").argtypes = [_I, _P, _P, _P, _P, _P, _P]
        getattr(L, f"orc_dot_{sfx}").argtypes = [_L, _P, _P, _P]
        getattr(L, f"orc_dot_{sfx}").restype = _D
        getattr(L, f"orc_true_residual_{sfx}").argtypes = [_I, _P, _P, _P, _P, _P, _P, _P]
        getattr(L, f"orc_true_residual_{sfx}").restype = _D
        getattr(L, f"orc_vcycle_{sfx}").argtypes = [_I, _P, _I, _I, _I, _P, _P, _P, _P, _P]
    L.orc_solve_f32.argtypes = [_I, _P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _D, _D, _D, _I,
                                _P, _I, _P]
    L.orc_solve_f64.argtypes = [_I, _P, _I, _I, _I, _P, _P, _P, _P, _P, _I, _D, _D, _D, _P, _I,
                                _P]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(shape):
    return (ctypes.c_int * 3)(*(list(shape) + [0] * (3 - len(shape))))


def _sfx(a):
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError(a.dtype)


def _check(a, dtype=None, shape=None):
    assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
    if dtype is not None:
        assert a.dtype == dtype, (a.dtype, dtype)
    if shape is not None:
        assert tuple(a.shape) == tuple(shape), (a.shape, shape)


def coarsen(fine):
    """init_gridtype (mgpcg.py:46-60)."""
    _check(fine, np.int8)
    dim = fine.ndim
    out = np.empty([s // 2 for s in fine.shape], np.int8)
    lib().orc_coarsen(dim, _dims(fine.shape), _ptr(fine), _ptr(out))
    return out


def smooth(gt, Adiag, Ax, r, z, phase):
    """smooth(l, phase) (mgpcg.py:104-109); updates z in place."""
    dim = gt.ndim
    for a in (Adiag, r, z):
        _check(a, Adiag.dtype, gt.shape)
    _check(Ax, Adiag.dtype, tuple(gt.shape) + (dim,))
    getattr(lib(), f"orc_smooth_{_sfx(Adiag)}")(dim, _dims(gt.shape), _ptr(gt), _ptr(Adiag),
                                                _ptr(Ax), _ptr(r), _ptr(z), int(phase))


def restrict(gt, Adiag, Ax, r, z):
    """restrict(l) (mgpcg.py:111-118) into a fresh zero-started coarse r."""
    dim = gt.ndim
    rc = np.zeros([s // 2 for s in gt.shape], Adiag.dtype)
    getattr(lib(), f"orc_restrict_{_sfx(Adiag)}")(dim, _dims(gt.shape), _ptr(gt), _ptr(Adiag),
                                                  _ptr(Ax), _ptr(r), _ptr(z), _ptr(rc))
    return rc


def prolongate(z, zc):
    """prolongate(l) (mgpcg.py:120-123); updates z in place."""
    _check(z)
    _check(zc, z.dtype, [s // 2 for s in z.shape])
    getattr(lib(), f"orc_prolongate_{_sfx(z)}")(z.ndim, _dims(z.shape), _ptr(z), _ptr(zc))


def compute_Ap(gt, Adiag, Ax, p, Ap=None):
    """compute_Ap (mgpcg.py:218-224); non-FLUID entries of Ap keep their value."""
    if Ap is None:
        Ap = np.zeros_like(p)
    getattr(lib(), f"orc_compute_Ap_{_sfx(Adiag)}")(gt.ndim, _dims(gt.shape), _ptr(gt),
                                                    _ptr(Adiag), _ptr(Ax), _ptr(p), _ptr(Ap))
    return Ap


def dot(gt, p, q):
    """reduce(p, q) (mgpcg.py:211-216), FP64 deterministic."""
    return getattr(lib(), f"orc_dot_{_sfx(p)}")(gt.size, _ptr(gt), _ptr(p), _ptr(q))


def true_residual(gt, Adiag, Ax, b, x):
    """r = b - A x in FP64; returns (r, r.r over FLUID)."""
    x = np.ascontiguousarray(x, np.float64)
    r = np.zeros(gt.shape, np.float64)
    rr = getattr(lib(), f"orc_true_residual_{_sfx(Adiag)}")(gt.ndim, _dims(gt.shape), _ptr(gt),
                                                            _ptr(Adiag), _ptr(Ax), _ptr(b),
                                                            _ptr(x), _ptr(r))
    return r, rr


def _ptr_array(arrs):
    arr = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arr


def v_cycle(h, r0, nu=2, bottom=10):
    """v_cycle (mgpcg.py:125-148) applied to r0; returns (z0, per-level r, z)."""
    L = len(h["gt"])
    dim = h["gt"][0].ndim
    dt = h["Adiag"][0].dtype
    r = [np.ascontiguousarray(r0, dt)] + [np.zeros_like(a) for a in h["Adiag"][1:]]
    z = [np.zeros_like(a) for a in h["Adiag"]]
    res = _dims(h["gt"][0].shape)
    getattr(lib(), f"orc_vcycle_{_sfx(h['Adiag'][0])}")(
        dim, res, L, nu, bottom, _ptr_array(h["gt"]), _ptr_array(h["Adiag"]),
        _ptr_array(h["Ax"]), _ptr_array(r), _ptr_array(z))
    return z[0], r, z


def solve(h, b, precision=PLAIN, nu=2, bottom=10, max_iters=500, rel_tol=1e-12, abs_tol=1e-14,
          eps=1e-12, refine_every=10, hist_cap=0):
    """solve (mgpcg.py:150-209) on hierarchy ``h`` = {gt, Adiag, Ax} lists.

    Returns (x as float64, result dict, rr history array).
    """
    L = len(h["gt"])
    dim = h["gt"][0].ndim
    res = _dims(h["gt"][0].shape)
    x = np.zeros(h["gt"][0].shape, np.float64)
    out = OrcResult()
    hist = np.zeros(max(hist_cap, 1), np.float64)
    for l in range(L):
        _check(h["gt"][l], np.int8)
        _check(h["Adiag"][l], h["Adiag"][0].dtype, h["gt"][l].shape)
        _check(h["Ax"][l], h["Adiag"][0].dtype, tuple(h["gt"][l].shape) + (dim,))
    if precision == F64:
        assert h["Adiag"][0].dtype == np.float64
        b = np.ascontiguousarray(b, np.float64)
        lib().orc_solve_f64(dim, res, L, nu, bottom, _ptr_array(h["gt"]), _ptr_array(h["Adiag"]),
                            _ptr_array(h["Ax"]), _ptr(b), _ptr(x), max_iters, rel_tol, abs_tol,
                            eps, _ptr(hist), hist_cap, ctypes.byref(out))
    else:
        assert h["Adiag"][0].dtype == np.float32
        b = np.ascontiguousarray(b, np.float32)
        lib().orc_solve_f32(dim, res, L, nu, bottom, int(precision == MIXED),
                            _ptr_array(h["gt"]), _ptr_array(h["Adiag"]), _ptr_array(h["Ax"]),
                            _ptr(b), _ptr(x), max_iters, rel_tol, abs_tol, eps, refine_every,
                            _ptr(hist), hist_cap, ctypes.byref(out))
    return x, out.as_dict(), hist[: min(hist_cap, out.iters + 1)]


def to_f64(h):
    """The same hierarchy with FP64 coefficients (for F64-mode solves)."""
    return {
        "gt": h["gt"],
        "Adiag": [np.ascontiguousarray(a, np.float64) for a in h["Adiag"]],
        "Ax": [np.ascontiguousarray(a, np.float64) for a in h["Ax"]],
    }
