"""C-ABI checks that need no GPU: the library loads and exports every entry
point declared in include/*.h, the host-side parameter code reproduces the
oracle bit for bit, triplet_to_col matches SciPy, and the UMFPACK status
codes come back for bad input (and device calls fail loudly without a GPU).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import germaine_amd as G
from germaine_amd import fdfd
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("gfdfd.h", "umfpack.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b((?:gfdfd|umfpack)_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = G.lib()
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
        assert name in G.SIGNATURES, f"{name} has no ctypes signature"


@pytest.mark.parametrize("s,free_surf", [(0.0, 0), (0.4, 1)])
def test_host_pml_profiles_bit_exact(s, free_surf):
    nx, ny, npml = 57, 33, 9
    for f in (2.0, 5.0, 16.0):
        prof = fdfd.pml_profiles(nx, ny, npml, 12.5, 90.0, f, s, free_surf)
        px, py = O.pml_profiles(nx, ny, npml, 12.5, 90.0, f, s, free_surf)
        want = np.concatenate([px.ravel(), py.ravel()])
        assert np.array_equal(prof.view(np.uint32), want.view(np.uint32))


def test_host_omega_and_source_bit_exact():
    for phys in (G.AC, G.SH, G.TE):
        for f, s in ((5.0, 0.0), (7.25, 0.3), (1.5e8, 0.0)):
            o2, om = fdfd.omega(phys, f, s)
            r2, rm = O.omega2(phys, f, s)
            assert np.array_equal(o2.view(np.uint32), r2.view(np.uint32))
            assert np.array_equal(om.view(np.uint32), rm.view(np.uint32))
            assert fdfd.source_amp(phys, f, s, 0.04) == O.source_amp(phys, f, s, 0.04)


def test_grid_index_matches_oracle():
    for x, y in ((500.0, 400.0), (29.9, 10.1), (12.5 * 37 + 6.2, 12.5 * 5)):
        assert fdfd.grid_index(x, y, 12.5, 30, 30, 1601) == O.source_index(x, y, 12.5, 30, 30, 1601)


def _i(a):
    return np.ascontiguousarray(a, np.int32)


def _d(a):
    return np.ascontiguousarray(a, np.float64)


@pytest.mark.parametrize("packed", [False, True])
def test_triplet_to_col_matches_scipy(packed):
    rng = np.random.default_rng(3)
    n, nz = 40, 300
    ti, tj = _i(rng.integers(0, n, nz)), _i(rng.integers(0, n, nz))
    tv = rng.standard_normal(nz) + 1j * rng.standard_normal(nz)
    tre, tim = _d(tv.real), _d(tv.imag)
    Ap, Ai = np.zeros(n + 1, np.int32), np.zeros(nz, np.int32)
    Ax, Az, Map = np.zeros(2 * nz), np.zeros(nz), np.zeros(nz, np.int32)
    if packed:
        tx = _d(np.stack([tv.real, tv.imag], 1).ravel())
        st = G.lib().umfpack_zi_triplet_to_col(n, n, nz, ti.ctypes.data, tj.ctypes.data, tx.ctypes.data,
                                               None, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, None,
                                               Map.ctypes.data)
        vals = Ax[0::2] + 1j * Ax[1::2]
    else:
        st = G.lib().umfpack_zi_triplet_to_col(n, n, nz, ti.ctypes.data, tj.ctypes.data,
                                               tre.ctypes.data, tim.ctypes.data, Ap.ctypes.data,
                                               Ai.ctypes.data, Ax.ctypes.data, Az.ctypes.data, Map.ctypes.data)
        vals = Ax[:nz] + 1j * Az
    assert st == 0
    S = sp.coo_matrix((tv, (ti, tj)), shape=(n, n)).tocsc()
    S.sum_duplicates()
    S.sort_indices()
    nnz = Ap[n]
    assert np.array_equal(Ap, S.indptr) and np.array_equal(Ai[:nnz], S.indices)
    np.testing.assert_allclose(vals[:nnz], S.data, rtol=1e-15, atol=1e-15)
    # Map: every triplet points at the entry it was summed into
    assert np.all(Ai[Map] == ti)
    col_of = np.repeat(np.arange(n), np.diff(Ap))
    assert np.all(col_of[Map] == tj)


def test_umfpack_status_codes_without_device():
    L = G.lib()
    sym = C.c_void_p()
    Ap = _i([0, 1, 2])
    Ai = _i([0, 1])
    Ax = _d([1.0, 2.0])
    info = np.zeros(90)
    assert L.umfpack_zi_symbolic(0, 0, Ap.ctypes.data, Ai.ctypes.data, None, None, C.byref(sym), None,
                                 info.ctypes.data) == -6
    assert info[0] == -6
    assert L.umfpack_zi_symbolic(2, 2, None, Ai.ctypes.data, None, None, C.byref(sym), None, None) == -5
    bad = _i([1, 0])                     # unsorted row indices
    Ap2 = _i([0, 2, 2])
    assert L.umfpack_zi_symbolic(2, 2, Ap2.ctypes.data, bad.ctypes.data, None, None, C.byref(sym), None, None) == -8
    assert L.umfpack_zi_symbolic(2, 3, Ap.ctypes.data, Ai.ctypes.data, None, None, C.byref(sym), None, None) == -8
    # the analysis itself is host-only
    assert L.umfpack_zi_symbolic(2, 2, Ap.ctypes.data, Ai.ctypes.data, None, None, C.byref(sym), None, None) == 0
    num = C.c_void_p()
    other = _i([0, 1])
    Ap3 = _i([0, 0, 2])
    assert L.umfpack_zi_numeric(Ap3.ctypes.data, other.ctypes.data, Ax.ctypes.data, None, sym, C.byref(num), None,
                                None) == -11
    assert L.umfpack_zi_numeric(Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, None, None, C.byref(num), None,
                                None) == -4
    x = np.zeros(4)
    assert L.umfpack_zi_solve(0, None, None, None, None, x.ctypes.data, None, x.ctypes.data, None, None, None,
                              None) == -3
    L.umfpack_zi_free_symbolic(C.byref(sym))
    assert sym.value is None
    L.umfpack_zi_free_symbolic(C.byref(sym))      # NULL-safe
    L.umfpack_zi_free_numeric(C.byref(num))


@pytest.mark.skipif(os.path.exists("/dev/kfd"),
                    reason="a GPU may be present")
def test_device_calls_fail_loudly_without_gpu():
    pl = fdfd.Plan(8, 6)
    vals = np.zeros(2 * pl.nnz)
    h = C.c_void_p()
    st = G.lib().gfdfd_factor(pl.h, 1, vals.ctypes.data, 64, None, C.byref(h))
    assert st == G.GFDFD_OK - 2          # GFDFD_ERR_DEVICE
    assert b"no HIP device" in G.lib().gfdfd_last_error()
