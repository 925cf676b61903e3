"""Pins for the oracle's assembly restatement (CPU only).

Each check is chosen so that a plausible slip in oracle/assemble.c -- a wrong
neighbour index, a dropped PML factor, a sign, the transpose -- fails it:

* an independent Kronecker-product construction of the Hustedt mixed-grid
  PML operator (rotated + standard Laplacian, lumped 9-point mass), built in
  float64 from 1-D difference operators, against the per-point triplets;
* for SH with varying mu, a float64 flux-form construction (face and corner
  fluxes with 2- and 4-point arithmetic averages of mu), with the reference's
  one departure from that form (its N entry) stated explicitly;
* the closed-form interior coefficients of Hustedt et al. (2004) for PML = 1;
* the closed-form nonzero count 9N - 6nx - 6ny + 4;
* a float64 closed form of the classical PML profile;
* libgcc's Smith division, against the C '/' operator on random operands;
* the CSC of M equals the transpose of the assembled stencil.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from tests._grids import smooth_model, te_model

A_W, B_W, D_W = 0.6287326, 0.3712667, 0.4382634


def f32(x):
    return float(np.float32(x))


def weights():
    a = f32(A_W)
    b = f32(B_W)
    c = np.float32(np.float32(1) - np.float32(a)) - np.float32(b)
    d = f32(D_W)
    e = f32(np.float32(1) - np.float32(d))
    return a, b, float(c), d, e


def second_diff(d0, dp, dm):
    """1-D PML second difference (f -> d0*(dp*(f+ - f) - dm*(f - f-))), zero outside."""
    n = len(d0)
    return sp.diags([d0[1:] * dm[1:], -d0 * (dp + dm), d0[:-1] * dp[:-1]], [-1, 0, 1], format="csr")


def kron_operator(nx, ny, dh, omega2, mass_col, blap, px, py):
    """float64 mixed-grid operator: row-point Laplacian weight blap, column mass."""
    a, b, c, d, e = weights()
    idh2 = 1.0 / (dh * dh)
    dx0 = px[0, 1:nx + 1] + 1j * px[1, 1:nx + 1]
    dxp = px[2, 1:nx + 1] + 1j * px[3, 1:nx + 1]
    dxm = px[2, 0:nx] + 1j * px[3, 0:nx]
    dy0 = py[0, 1:ny + 1] + 1j * py[1, 1:ny + 1]
    dyp = py[2, 1:ny + 1] + 1j * py[3, 1:ny + 1]
    dym = py[2, 0:ny] + 1j * py[3, 0:ny]
    Dxx = second_diff(dx0, dxp, dxm)
    Dyy = second_diff(dy0, dyp, dym)
    Tx = sp.diags([np.ones(nx - 1), np.ones(nx - 1)], [-1, 1])
    Ty = sp.diags([np.ones(ny - 1), np.ones(ny - 1)], [-1, 1])
    Ix, Iy = sp.identity(nx), sp.identity(ny)
    Sx, Sy = Tx + 2 * Ix, Ty + 2 * Iy
    Lrot = 0.25 * idh2 * (sp.kron(Dyy, Sx) + sp.kron(Sy, Dxx))
    Lstd = idh2 * (sp.kron(Dyy, Ix) + sp.kron(Iy, Dxx))
    W = a * sp.identity(nx * ny) + 0.25 * b * (sp.kron(Iy, Tx) + sp.kron(Ty, Ix)) + 0.25 * c * sp.kron(Ty, Tx)
    A = omega2 * W @ sp.diags(mass_col.ravel()) + sp.diags(blap.ravel()) @ (d * Lrot + e * Lstd)
    return A.tocsr()


def triplets_to_csr(nx, ny, irow, icol, ar, ai):
    n = nx * ny
    return sp.csr_matrix((ar + 1j * ai, (irow, icol)), shape=(n, n))


def assert_rowwise_close(A, B, rtol):
    D = (A - B).tocsr()
    scale = abs(B).max(axis=1).toarray().ravel()
    err = abs(D).max(axis=1).toarray().ravel()
    assert np.all(err <= rtol * scale), f"max rel {np.max(err / scale):.3e}"


@pytest.mark.parametrize("nx,ny,npml", [(5, 4, 0), (12, 9, 3), (40, 30, 8)])
def test_nonzero_closed_form(nx, ny, npml):
    assert O.nonzero(nx, ny) == 9 * nx * ny - 6 * nx - 6 * ny + 4


def test_nonzero_configs():
    # grid sizes of the five workloads (full grid incl. PML)
    for nx, ny, nnz in [(201, 101, 180901), (801, 187, 1342159), (1601, 401, 5766001),
                        (801, 201, 1443001), (2401, 801, 17289601)]:
        assert 9 * nx * ny - 6 * nx - 6 * ny + 4 == nnz
    assert O.nonzero(201, 101) == 180901


@pytest.mark.parametrize("s", [0.0, 0.7])
@pytest.mark.parametrize("free_surf", [0, 1])
def test_pml_profile_closed_form(s, free_surf):
    nx, ny, npml, dh, a0, f = 40, 30, 8, 10.0, 90.0, 5.0
    px, py = O.pml_profiles(nx, ny, npml, dh, a0, f, s, free_surf)
    L = npml * dh
    om = 2 * np.pi * (f - 1j * s)
    ref = np.ones(nx + 2, complex)
    refh = np.ones(nx + 2, complex)
    for i in range(1, npml + 1):
        x = i * dh
        for arr, xx in ((ref, x), (refh, x + 0.5 * dh)):
            eps = a0 * (1 - np.cos((L - xx) * np.pi / 2 / L))
            arr[i] = arr[nx - i + 1] = 1 / (1 + eps / (1j * om))
    ref[0], ref[-1], refh[0], refh[-1] = ref[1], ref[nx], refh[1], refh[nx]
    np.testing.assert_allclose(px[0] + 1j * px[1], ref, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(px[2] + 1j * px[3], refh, rtol=2e-6, atol=1e-7)
    # y profile: same values; the top strip is switched off with a free surface
    y = py[0] + 1j * py[1]
    np.testing.assert_allclose(y[ny - npml + 1:ny + 2], (px[0] + 1j * px[1])[nx - npml + 1:nx + 2], rtol=0)
    if free_surf:
        assert np.all(y[0:npml + 1] == 1.0)
    else:
        np.testing.assert_allclose(y[0:npml + 1], (px[0] + 1j * px[1])[0:npml + 1], rtol=0)
    # the inner edge of the PML is undamped
    assert px[0, npml] == 1.0 and px[1, npml] == 0.0


def test_omega2_matches_square():
    for phys, sgn in ((O.AC, -1), (O.SH, 1), (O.TE, 1)):
        for f, s in ((5.0, 0.0), (12.5, 0.3), (3.0e8, 0.0)):
            o2, om = O.omega2(phys, f, s)
            w = 2 * np.pi * (f + sgn * 1j * s)
            np.testing.assert_allclose(o2[0] + 1j * o2[1], w * w, rtol=3e-7, atol=3e-7 * abs(w * w))
            np.testing.assert_allclose(om[0] + 1j * om[1], w, rtol=1e-7)


@pytest.mark.parametrize("s", [0.0, 0.5])
def test_ac_matches_kronecker(s):
    nx, ny, npml, dh, f = 36, 27, 7, 12.5, 6.0
    vp, rho = smooth_model(nx, ny, seed=3)
    px, py = O.pml_profiles(nx, ny, npml, dh, 90.0, f, s, 0)
    tri = O.assemble(O.AC, nx, ny, dh, f, s, vp, rho, px, py)
    A = triplets_to_csr(nx, ny, *tri)
    o2, _ = O.omega2(O.AC, f, s)
    ivp2 = 1.0 / (rho.astype(np.float64) * vp * vp)
    K = kron_operator(nx, ny, dh, o2[0] + 1j * o2[1], ivp2, 1.0 / rho.astype(np.float64), px, py)
    assert_rowwise_close(A, K, 2e-6)
    # the pattern is exactly the 9-point one, entries in ascending column order per row
    irow, icol = tri[0], tri[1]
    assert np.all(np.diff(irow) >= 0)
    same = irow[1:] == irow[:-1]
    assert np.all(np.diff(icol)[same] > 0)


def test_te_matches_kronecker():
    nx, ny, npml, dh, f, s = 30, 24, 6, 0.04, 250e6, 0.0
    eps, sig = te_model(nx, ny, seed=5)
    px, py = O.pml_profiles(nx, ny, npml, dh, 90.0, f, s, 0)
    A = triplets_to_csr(nx, ny, *O.assemble(O.TE, nx, ny, dh, f, s, eps, sig, px, py))
    o2, om = O.omega2(O.TE, f, s)
    w = om[0] + 1j * om[1]
    mu0 = f32(4 * np.pi * 1e-7)
    mass = mu0 * (eps.astype(np.float64) + 1j * sig / w)
    K = kron_operator(nx, ny, dh, o2[0] + 1j * o2[1], mass, np.ones((ny, nx)), px, py)
    assert_rowwise_close(A, K, 2e-6)


def test_sh_constant_mu_matches_kronecker():
    # SH with constant mu: all nine averaged moduli equal mu (src/SH/av_mu.c:12-28)
    nx, ny, npml, dh, f = 30, 22, 6, 10.0, 8.0
    vs = np.full((ny, nx), 1500.0, np.float32)
    rng = np.random.default_rng(1)
    rho = (1800 + 200 * rng.random((ny, nx))).astype(np.float32)
    rho[:] = 2000.0
    px, py = O.pml_profiles(nx, ny, npml, dh, 90.0, f, 0.0, 0)
    A = triplets_to_csr(nx, ny, *O.assemble(O.SH, nx, ny, dh, f, 0.0, vs, rho, px, py))
    o2, _ = O.omega2(O.SH, f, 0.0)
    mu = np.float32(np.float32(2000.0) * np.float32(1500.0)) * np.float32(1500.0)
    K = kron_operator(nx, ny, dh, o2[0] + 1j * o2[1], rho.astype(np.float64), np.full((ny, nx), float(mu)), px, py)
    assert_rowwise_close(A, K, 2e-6)


def sh_flux_operator(nx, ny, dh, omega2, rho, mu, px, py, n_uses_bm0=True):
    """float64 variable-coefficient SH operator in flux form, built per grid point.

    Standard Laplacian: face fluxes with the 2-point arithmetic averages of mu
    at (i+-1/2, j) and (i, j+-1/2).  Rotated Laplacian: one flux per cell
    corner (i+-1/2, j+-1/2) with the 4-point average of mu there, its x part
    dx0*dx_c*mu_c*[(u(i+sx,j) - u(i,j)) + (u(i+sx,j+sy) - u(i,j+sy))]/(4h^2) and
    its y part the same with the axes exchanged.  mu outside the grid is the
    nearest edge value (src/SH/init_mat_SH.c:28-32).  Entries whose column
    falls outside the grid are dropped; the centre keeps its whole sum.

    n_uses_bm0: the reference's N entry multiplies its standard-Laplacian
    flux by the x average mu(i-1/2, j) rather than mu(i, j-1/2)
    (src/SH/init_A_SH_9p_pml.c:94, `(*matSH).bm0 * dym`), while the centre
    uses mu(i, j-1/2) (:144).  The flag states that one departure from the
    flux form explicitly; every other coefficient follows the mathematics.
    """
    a, b, c, d, e = weights()
    h2i = 1.0 / (dh * dh)
    mu = mu.astype(np.float64)
    M = np.empty((ny + 2, nx + 2))
    M[1:-1, 1:-1] = mu
    M[1:-1, 0], M[1:-1, -1] = mu[:, 0], mu[:, -1]
    M[0, :], M[-1, :] = M[1, :], M[-2, :]
    x0, xh = px[0] + 1j * px[1], px[2] + 1j * px[3]   # [0..nx+1]
    y0, yh = py[0] + 1j * py[1], py[2] + 1j * py[3]
    rows, cols, vals = [], [], []
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            k = (i - 1) + (j - 1) * nx
            ent = {}

            def add(ii, jj, v):
                ent[(ii, jj)] = ent.get((ii, jj), 0.0) + v

            # standard Laplacian (weight e)
            fx = {+1: 0.5 * (M[j, i] + M[j, i + 1]) * xh[i], -1: 0.5 * (M[j, i] + M[j, i - 1]) * xh[i - 1]}
            fy = {+1: 0.5 * (M[j, i] + M[j + 1, i]) * yh[j], -1: 0.5 * (M[j, i] + M[j - 1, i]) * yh[j - 1]}
            for s in (+1, -1):
                add(i + s, j, e * h2i * x0[i] * fx[s])
                add(i, j, -e * h2i * x0[i] * fx[s])
                coef = fy[s]
                if s == -1 and n_uses_bm0:
                    coef = 0.5 * (M[j, i] + M[j, i - 1]) * yh[j - 1]
                add(i, j + s, e * h2i * y0[j] * coef)
                add(i, j, -e * h2i * y0[j] * fy[s])
            # rotated Laplacian (weight d): one flux per corner
            for sx in (+1, -1):
                for sy in (+1, -1):
                    mc = 0.25 * (M[j, i] + M[j, i + sx] + M[j + sy, i] + M[j + sy, i + sx])
                    cx_ = 0.25 * h2i * x0[i] * (xh[i] if sx > 0 else xh[i - 1]) * mc
                    cy_ = 0.25 * h2i * y0[j] * (yh[j] if sy > 0 else yh[j - 1]) * mc
                    for (ii, jj), sg in (((i + sx, j), 1), ((i, j), -1), ((i + sx, j + sy), 1), ((i, j + sy), -1)):
                        add(ii, jj, d * cx_ * sg)
                    for (ii, jj), sg in (((i, j + sy), 1), ((i, j), -1), ((i + sx, j + sy), 1), ((i + sx, j), -1)):
                        add(ii, jj, d * cy_ * sg)
            # lumped mass of the column point
            wts = {0: a, 1: 0.25 * b, 2: 0.25 * c}
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    ii, jj = i + di, j + dj
                    if 1 <= ii <= nx and 1 <= jj <= ny:
                        add(ii, jj, wts[abs(di) + abs(dj)] * omega2 * float(rho[jj - 1, ii - 1]))
            for (ii, jj), v in ent.items():
                if 1 <= ii <= nx and 1 <= jj <= ny:
                    rows.append(k)
                    cols.append((ii - 1) + (jj - 1) * nx)
                    vals.append(v)
    n = nx * ny
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


@pytest.mark.parametrize("s", [0.0, 0.4])
def test_sh_varying_mu_matches_flux_form(s):
    """SH with random rho and vs (so mu varies cell to cell) against the flux-form
    construction: pins the 4-point corner and 2-point face averages of
    src/SH/av_mu.c:12-28 and their use in src/SH/init_A_SH_9p_pml.c:68-213."""
    nx, ny, npml, dh, f = 26, 19, 5, 10.0, 7.0
    rng = np.random.default_rng(11)
    vs = (1000 + 1500 * rng.random((ny, nx))).astype(np.float32)
    rho = (1500 + 1000 * rng.random((ny, nx))).astype(np.float32)
    px, py = O.pml_profiles(nx, ny, npml, dh, 90.0, f, s, 0)
    A = triplets_to_csr(nx, ny, *O.assemble(O.SH, nx, ny, dh, f, s, vs, rho, px, py))
    o2, _ = O.omega2(O.SH, f, s)
    mu = (rho * vs) * vs   # float, as src/SH/init_mat_SH.c:22
    K = sh_flux_operator(nx, ny, dh, o2[0] + 1j * o2[1], rho, mu, px, py)
    assert_rowwise_close(A, K, 2e-6)
    # without the N-entry departure the forms differ exactly in the N entries,
    # so a wrong neighbour in any average would show
    K0 = sh_flux_operator(nx, ny, dh, o2[0] + 1j * o2[1], rho, mu, px, py, n_uses_bm0=False)
    D = abs(A - K0).tocoo()
    big = D.data > 1e-4 * abs(A).max()
    assert big.any()
    assert np.all(D.col[big] == D.row[big] - nx)


def test_sh_flux_form_detects_swapped_corner():
    """A corner average taken at the wrong corner (bpm <-> bmp) in the oracle's
    mathematics would fail the comparison above: the flux form itself changes."""
    nx, ny, dh, f = 12, 10, 10.0, 6.0
    rng = np.random.default_rng(3)
    vs = (1000 + 1500 * rng.random((ny, nx))).astype(np.float32)
    rho = np.full((ny, nx), 2000.0, np.float32)
    px, py = O.pml_profiles(nx, ny, 0, dh, 90.0, f, 0.0, 0)
    A = triplets_to_csr(nx, ny, *O.assemble(O.SH, nx, ny, dh, f, 0.0, vs, rho, px, py))
    o2, _ = O.omega2(O.SH, f, 0.0)
    mu = (rho * vs) * vs
    # transposed mu field: the averages move to other corners, the operator changes
    K = sh_flux_operator(nx, ny, dh, o2[0] + 1j * o2[1], rho, mu[::-1, ::-1].copy(), px, py)
    with pytest.raises(AssertionError):
        assert_rowwise_close(A, K, 2e-6)


def test_interior_hustedt_coefficients():
    """PML = 1 and constant material: the textbook mixed-grid stencil."""
    nx, ny, dh, f = 9, 8, 10.0, 4.0
    vp = np.full((ny, nx), 2000.0, np.float32)
    rho = np.full((ny, nx), 1000.0, np.float32)
    px, py = O.pml_profiles(nx, ny, 0, dh, 90.0, f, 0.0, 0)
    A = triplets_to_csr(nx, ny, *O.assemble(O.AC, nx, ny, dh, f, 0.0, vp, rho, px, py)).toarray()
    a, b, c, d, e = weights()
    w2 = (2 * np.pi * f) ** 2
    m = w2 / (1000.0 * 2000.0 ** 2)
    bb = 1.0 / 1000.0
    h2 = dh * dh
    k = 4 + 4 * nx
    corner = c / 4 * m + d * bb / (2 * h2)
    edge = b / 4 * m + e * bb / h2
    centre = a * m - (2 * d + 4 * e) * bb / h2
    for col, want in ((k - 1 - nx, corner), (k - nx, edge), (k + 1 - nx, corner), (k - 1, edge), (k, centre),
                      (k + 1, edge), (k - 1 + nx, corner), (k + nx, edge), (k + 1 + nx, corner)):
        assert abs(A[k, col] - want) <= 2e-6 * abs(want)
    # consistency: the Laplacian part annihilates constants, so the row sum is the mass
    assert abs(A[k].sum() - m) <= 1e-6 * abs(centre)


def test_csc_of_M_is_transpose():
    nx, ny = 7, 5
    vp, rho = smooth_model(nx, ny, seed=2)
    px, py = O.pml_profiles(nx, ny, 2, 10.0, 90.0, 5.0, 0.0, 0)
    tri = O.assemble(O.AC, nx, ny, 10.0, 5.0, 0.0, vp, rho, px, py)
    A = triplets_to_csr(nx, ny, *tri)
    ap, ai, ax, az = O.csc_of_M(nx, ny, *tri)
    M = sp.csc_matrix((ax + 1j * az, ai, ap), shape=A.shape)
    assert abs(M - A.T).max() == 0
    assert abs(M - A).max() > 0          # A is not symmetric (mass and PML terms)


def test_smith_division_matches_libgcc():
    rng = np.random.default_rng(7)
    vals = (rng.standard_normal((20000, 4)) * 10.0 ** rng.integers(-12, 12, (20000, 4))).astype(np.float32)
    vals[::3, 1] = 0.0
    import ctypes as C
    L = O.lib()
    x1, y1, x2, y2 = C.c_float(), C.c_float(), C.c_float(), C.c_float()
    for a, b, c, d in vals:
        L.orc_smith_divf(a, b, c, d, C.byref(x1), C.byref(y1))
        L.orc_c_divf(a, b, c, d, C.byref(x2), C.byref(y2))
        assert (x1.value, y1.value) == (x2.value, y2.value)


def test_source_index_known_answers():
    # row k = nxsrc + (nysrc-1)*NX with nxsrc = iround(x/DH)+NPML, nysrc = iround(y/DH)+FSSHIFT
    # (src/AC/RHS_source_AC.c:23-32); hand-computed for a 270-wide grid, DH=20, NPML=10
    assert O.source_index(500.0, 400.0, 20.0, 10, 10, 270) == 35 + 29 * 270
    assert O.source_index(9500.0, 400.0, 20.0, 10, 10, 270) == 485 + 29 * 270
    assert O.source_index(29.9, 10.1, 20.0, 0, 0, 50) == 1 + 0 * 50      # iround(1.495)=1, iround(0.505)=1
    # amp/(DH*DH) is a float division stored into double
    assert O.source_amp(O.AC, 5.0, 0.0, 20.0) == complex(float(np.float32(1) / np.float32(400)), 0.0)
    # TE: -(i omega mu0)/DH^2 = (omega mu0 (imag), -omega mu0 (real)) / DH^2 for S=0
    f, dh = 1e8, 0.04
    amp = O.source_amp(O.TE, f, 0.0, dh)
    w = 2 * np.pi * f * 4e-7 * np.pi
    assert abs(amp.imag + w / dh ** 2) <= 1e-6 * w / dh ** 2 and amp.real == 0.0
