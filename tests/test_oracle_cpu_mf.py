"""Pins for the CPU FP64 multifrontal oracle (oracle/mf_cpu.c), CPU only.

* dense LAPACK solves of M and M^T on small AC, AC with S > 0 and TE grids;
* random complex values on the 9-point pattern, which force in-front row
  swaps (catches pivot bookkeeping in the transposed solve);
* the banded-LU oracle on the 201x101 acoustic grid with ten point sources;
* threads do not change the answer (the tree levels are independent).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import mf
from oracle import oracle as O
from tests.test_oracle_solver import ac_system, dense, te_system


@pytest.mark.parametrize("sysno", [0, 2])
@pytest.mark.parametrize("kind,leaf", [("ac", 8), ("ac_s", 32), ("te", 6)])
def test_mf_vs_dense(kind, leaf, sysno):
    nx, ny = 23, 17
    csc = {"ac": lambda: ac_system(nx, ny, 5, 10.0, 12.0),
           "ac_s": lambda: ac_system(nx, ny, 5, 10.0, 12.0, s=0.8, seed=4),
           "te": lambda: te_system(nx, ny, 5, 0.04, 300e6)}[kind]()
    M = dense(*csc)
    rng = np.random.default_rng(sysno)
    B = rng.standard_normal((3, nx * ny)) + 1j * rng.standard_normal((3, nx * ny))
    F = mf.Factor(nx, ny, *csc, leaf=leaf)
    X = F.solve(B, sysno)
    Xd = np.linalg.solve(M if sysno == 0 else M.T, B.T).T
    assert np.max(np.abs(X - Xd)) <= 1e-12 * np.max(np.abs(Xd))


@pytest.mark.parametrize("sysno", [0, 2])
def test_mf_random_pattern_pivoting(sysno):
    nx, ny = 24, 19
    n = nx * ny
    M0 = sp.csc_matrix(dense(*ac_system(nx, ny, 4, 10.0, 9.0)))
    rng = np.random.default_rng(11 + sysno)
    v = rng.standard_normal(M0.nnz) + 1j * rng.standard_normal(M0.nnz)
    ap, ai = M0.indptr.astype(np.int32), M0.indices.astype(np.int32)
    M = sp.csc_matrix((v, ai, ap), shape=(n, n)).toarray()
    F = mf.Factor(nx, ny, ap, ai, v.real.copy(), v.imag.copy(), leaf=16)
    B = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    X = F.solve(B, sysno)
    Xd = np.linalg.solve(M if sysno == 0 else M.T, B.T).T
    assert np.max(np.abs(X - Xd)) <= 1e-9 * np.max(np.abs(Xd))


def test_mf_vs_band_config_grid():
    """201x101 grid (20-cell PML, h=10 m, 5 Hz), ten point sources, M and M^T;
    thread count does not change the result."""
    nx, ny = 201, 101
    csc = ac_system(nx, ny, 20, 10.0, 5.0, seed=1)
    n = nx * ny
    B = np.zeros((10, n), complex)
    B[np.arange(10), 5 * nx + 20 + 16 * np.arange(10)] = 1.0
    for sysno in (0, 2):
        Xb, info = O.band_solve(*csc, B.T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        assert info == 0
        mf.set_threads(1)
        X1 = mf.Factor(nx, ny, *csc).solve(B, sysno)
        mf.set_threads(4)
        X4 = mf.Factor(nx, ny, *csc).solve(B, sysno)
        assert np.max(np.abs(X1 - Xb.T)) <= 1e-11 * np.max(np.abs(Xb))
        assert np.max(np.abs(X4 - X1)) <= 1e-13 * np.max(np.abs(X1))
