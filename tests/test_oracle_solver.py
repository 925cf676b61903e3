"""Pins for the oracle's FP64 banded-LU solver (CPU only).

* numpy/LAPACK dense solves of M, M^H and M^T on small assembled grids
  (AC and TE, S = 0 and S > 0) -- catches wrong transposes, conjugations,
  pivot bookkeeping and band indexing;
* SciPy SuperLU on the 201x101 acoustic grid with multiple right-hand sides;
* the oracle-free componentwise backward error.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from tests._grids import smooth_model, te_model


def ac_system(nx, ny, npml, dh, f, s=0.0, seed=0):
    vp, rho = smooth_model(nx, ny, seed=seed)
    px, py = O.pml_profiles(nx, ny, npml, dh, 90.0, f, s, 0)
    return O.csc_of_M(nx, ny, *O.assemble(O.AC, nx, ny, dh, f, s, vp, rho, px, py))


def te_system(nx, ny, npml, dh, f, seed=0):
    eps, sig = te_model(nx, ny, seed=seed)
    px, py = O.pml_profiles(nx, ny, npml, dh, 90.0, f, 0.0, 0)
    return O.csc_of_M(nx, ny, *O.assemble(O.TE, nx, ny, dh, f, 0.0, eps, sig, px, py))


def dense(ap, ai, ax, az):
    n = len(ap) - 1
    return sp.csc_matrix((ax + 1j * az, ai, ap), shape=(n, n)).toarray()


@pytest.mark.parametrize("sysno", [0, 1, 2])
@pytest.mark.parametrize("kind", ["ac", "ac_s", "te"])
def test_band_vs_dense(sysno, kind):
    nx, ny = 23, 17
    if kind == "ac":
        csc = ac_system(nx, ny, 5, 10.0, 12.0)
    elif kind == "ac_s":
        csc = ac_system(nx, ny, 5, 10.0, 12.0, s=0.8, seed=4)
    else:
        csc = te_system(nx, ny, 5, 0.04, 300e6)
    M = dense(*csc)
    op = {0: M, 1: M.conj().T, 2: M.T}[sysno]
    rng = np.random.default_rng(sysno)
    B = rng.standard_normal((nx * ny, 3)) + 1j * rng.standard_normal((nx * ny, 3))
    X, info = O.band_solve(*csc, B, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
    assert info == 0
    Xd = np.linalg.solve(op, B)
    assert np.max(np.abs(X - Xd)) <= 1e-12 * np.max(np.abs(Xd))
    X2, _ = O.band_solve(*csc, B, sys=sysno)          # no permutation: same answer
    assert np.max(np.abs(X2 - Xd)) <= 1e-12 * np.max(np.abs(Xd))


def test_band_vs_superlu_config_grid():
    """201x101 grid (20-cell PML, h=10 m, 5 Hz): ten point sources."""
    nx, ny = 201, 101
    csc = ac_system(nx, ny, 20, 10.0, 5.0, seed=1)
    n = nx * ny
    B = np.zeros((n, 10), complex)
    for s in range(10):
        B[(20 + 16 * s) + 24 * nx, s] = 1.0 / 100.0
    X, info = O.band_solve(*csc, B, sys=0, perm=O.grid_transpose_perm(nx, ny))
    assert info == 0
    M = sp.csc_matrix((csc[2] + 1j * csc[3], csc[1], csc[0]), shape=(n, n))
    Xs = spla.splu(M).solve(B)
    err = np.linalg.norm(X - Xs, axis=0) / np.linalg.norm(Xs, axis=0)
    assert np.all(err <= 1e-10), err
    # oracle-free componentwise backward error
    R = B - M @ X
    den = abs(M) @ np.abs(X) + np.abs(B)
    assert np.max(np.abs(R) / den) <= 1e-13


def test_stencil_diagonals_match_oracle_matvec():
    """The nine-diagonal test operator (tests/_stencil.py) equals the oracle's
    CSC matvec of M and M^T on a PML grid with S > 0."""
    from tests import _stencil as ST
    from tests._grids import smooth_model
    nx, ny = 23, 17
    vp, rho = smooth_model(nx, ny, seed=4)
    px, py = O.pml_profiles(nx, ny, 4, 10.0, 90.0, 6.0, 0.3, 0)
    tri = O.assemble(O.AC, nx, ny, 10.0, 6.0, 0.3, vp, rho, px, py)
    csc = O.csc_of_M(nx, ny, *tri)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((3, nx * ny)) + 1j * rng.standard_normal((3, nx * ny))
    for sysno, tr in ((0, False), (2, True)):
        D = ST.diagonals(nx * ny, *tri, transpose=tr)
        assert len(D) == 9
        Y = ST.apply(D, X)
        for c in range(3):
            ref = O.matvec(*csc, X[c], sysno)
            assert np.max(np.abs(Y[c] - ref)) <= 1e-12 * np.max(np.abs(ref))
    B = ST.apply(ST.diagonals(nx * ny, *tri), X)
    assert np.max(ST.relres(ST.diagonals(nx * ny, *tri), X, B)) <= 1e-15
