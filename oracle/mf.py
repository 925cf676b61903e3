"""ORACLE — test infrastructure only.

ctypes front-end of ``libmfcpu.so`` (oracle/mf_cpu.c): an FP64 multifrontal
LU of M with OpenMP on the host cores, used as an FP64 checker on grids too
large for the banded oracle and as the timed CPU baseline.  Only ``tests/``
and ``bench.py``'s ``cpu_baseline`` leg import it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libmfcpu.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        L = C.CDLL(path)
        f = np.ctypeslib.ndpointer
        L.mfc_factor.restype = C.c_void_p
        L.mfc_factor.argtypes = [C.c_int, C.c_int, f(np.int32), f(np.int32), f(np.float64), f(np.float64),
                                 C.c_int, C.POINTER(C.c_int)]
        L.mfc_solve.restype = None
        L.mfc_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, f(np.float64), f(np.float64),
                                f(np.float64), f(np.float64)]
        L.mfc_factor_flops.restype = C.c_double
        L.mfc_factor_flops.argtypes = [C.c_void_p]
        L.mfc_threads.restype = C.c_int
        L.mfc_set_threads.argtypes = [C.c_int]
        L.mfc_free.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def set_threads(t):
    if t:
        lib().mfc_set_threads(int(t))
    return lib().mfc_threads()


class Factor:
    """FP64 multifrontal factors of the grid matrix given in CSC (Ap, Ai, Ax, Az)."""

    def __init__(self, nx, ny, ap, ai, ax, az, leaf=32):
        self.n = nx * ny
        sing = C.c_int(0)
        self._keep = [np.ascontiguousarray(a, t) for a, t in
                      ((ap, np.int32), (ai, np.int32), (ax, np.float64), (az, np.float64))]
        self.h = lib().mfc_factor(nx, ny, *self._keep, leaf, C.byref(sing))
        self.singular = bool(sing.value)

    @property
    def flops(self):
        return lib().mfc_factor_flops(self.h)

    def solve(self, b, sys=0):
        """op(M) x = b for sys 0 (M) or 2 (M^T); b complex [n] or [nrhs, n]."""
        assert sys in (0, 2)
        b = np.asarray(b)
        one = b.ndim == 1
        B = b.reshape(-1, self.n)
        nrhs = B.shape[0]
        br = np.ascontiguousarray(B.real, np.float64).ravel()
        bi = np.ascontiguousarray(B.imag, np.float64).ravel()
        xr = np.empty_like(br)
        xi = np.empty_like(bi)
        lib().mfc_solve(self.h, sys, nrhs, br, bi, xr, xi)
        x = (xr + 1j * xi).reshape(nrhs, self.n)
        return x[0] if one else x

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.mfc_free(self.h)
            self.h = None
