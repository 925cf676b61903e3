"""GPU parity of the HIP path against the oracle, through the C-ABI.

* device assembly: bit-identical to the oracle's triplet values (AC, SH, TE;
  S = 0 and S > 0; free surface) -- integer-exact comparison of the doubles;
* FP64 factor + solve: M, M^H, M^T against the FP64 banded-LU oracle at
  relative error <= 1e-10, through both the small-front (LDS) and the blocked
  (panel/TRSM/MFMA) paths, 1 to 70 right-hand sides;
* FP32 factors + FP64 refinement: <= 1e-8 against the oracle;
* batched frequencies equal per-frequency solves;
* every solve-kernel path -- single-block levels (one product per front),
  multi-block levels on the fused per-front kernels at each tile width, and
  on the per-block kernels -- for M, M^H, M^T in FP64 and FP32+IR, on
  assembled and on random-valued (row-swapping) matrices;
* refactoring in place equals a fresh factorisation;
* the 201x101 acoustic workload through the six umfpack_zi_* entry points.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import germaine_amd as G
from germaine_amd import fdfd, models
from oracle import oracle as O
from tests._grids import smooth_model, te_model

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def rel(x, ref):
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def oracle_system(phys, nx, ny, npml, dh, f, m1, m2, s=0.0, fs=0):
    px, py = O.pml_profiles(nx, ny, npml, dh, 90.0, f, s, fs)
    tri = O.assemble(phys, nx, ny, dh, f, s, m1, m2, px, py)
    return tri, O.csc_of_M(nx, ny, *tri)


@pytest.mark.parametrize("phys,s,fs", [(G.AC, 0.0, 0), (G.AC, 0.5, 1), (G.SH, 0.3, 0), (G.TE, 0.0, 0)])
def test_assembly_bit_exact(phys, s, fs):
    nx, ny, npml = 37, 29, 6
    if phys == G.TE:
        m1, m2 = te_model(nx, ny, seed=2)
        dh, freqs = 0.04, [1.0e8, 3.5e8]
    else:
        m1, m2 = smooth_model(nx, ny, seed=5)
        dh, freqs = 10.0, [4.0, 11.5]
    say("assembly", phys, s, fs)
    pl = fdfd.Plan(nx, ny)
    vals = pl.assemble(phys, dh, freqs, dev(m1), dev(m2), s=s, npml=npml, free_surf=fs).cpu().numpy()
    for k, f in enumerate(freqs):
        (irow, icol, ar, ai), _ = oracle_system(phys, nx, ny, npml, dh, f, m1, m2, s, fs)
        assert np.array_equal(vals[k].real, ar), f"real parts differ at {np.nonzero(vals[k].real != ar)[0][:5]}"
        assert np.array_equal(vals[k].imag, ai), f"imag parts differ at {np.nonzero(vals[k].imag != ai)[0][:5]}"


def system(nx=45, ny=33, npml=6, f=9.0, seed=1):
    m1, m2 = smooth_model(nx, ny, seed=seed)
    tri, csc = oracle_system(G.AC, nx, ny, npml, 10.0, f, m1, m2)
    return m1, m2, csc


@pytest.mark.parametrize("small_max", [64, 16])
@pytest.mark.parametrize("sysno", [0, 1, 2])
@pytest.mark.parametrize("nrhs", [1, 3, 70])
def test_fp64_solve_matches_oracle(small_max, sysno, nrhs):
    nx, ny = 45, 33
    m1, m2, csc = system(nx, ny)
    say("fp64 solve", small_max, sysno, nrhs)
    pl = fdfd.Plan(nx, ny, leaf_area=32, small_max=small_max)
    vals = dev((csc[2] + 1j * csc[3])[None, :])
    F = pl.factor(vals, G.FP64)
    assert F.singular()[0] == 0
    rng = np.random.default_rng(nrhs)
    B = rng.standard_normal((nrhs, nx * ny)) + 1j * rng.standard_normal((nrhs, nx * ny))
    X = F.solve(0, dev(B), sys=sysno).cpu().numpy()
    Xr, info = O.band_solve(*csc, B.T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
    assert rel(X, Xr.T) <= 1e-10


@pytest.mark.parametrize("sysno", [0, 2])
def test_fp32_factors_with_refinement(sysno):
    nx, ny = 45, 33
    m1, m2, csc = system(nx, ny)
    say("fp32+IR", sysno)
    pl = fdfd.Plan(nx, ny, small_max=16)
    F = pl.factor(dev((csc[2] + 1j * csc[3])[None, :]), G.FP32)
    rng = np.random.default_rng(7)
    B = rng.standard_normal((5, nx * ny)) + 1j * rng.standard_normal((5, nx * ny))
    X, it, rr = F.solve_refine(0, dev(B), sys=sysno, tol=1e-12, max_iter=20)
    X = X.cpu().numpy()
    Xr, _ = O.band_solve(*csc, B.T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
    say("  iterations", it, "relres", rr)
    assert rr <= 1e-12 and it >= 1
    assert rel(X, Xr.T) <= 1e-8


def test_batched_frequencies_equal_single():
    nx, ny, npml, dh = 41, 30, 6, 10.0
    m1, m2 = smooth_model(nx, ny, seed=9)
    freqs = [3.0, 6.5, 12.0]
    say("batched frequencies")
    pl = fdfd.Plan(nx, ny, small_max=16)
    vals = pl.assemble(G.AC, dh, freqs, dev(m1), dev(m2), npml=npml)
    F = pl.factor(vals, G.FP64)
    rng = np.random.default_rng(3)
    B = rng.standard_normal((4, nx * ny)) + 1j * rng.standard_normal((4, nx * ny))
    for k, f in enumerate(freqs):
        _, csc = oracle_system(G.AC, nx, ny, npml, dh, f, m1, m2)
        X = F.solve(k, dev(B)).cpu().numpy()
        Xr, _ = O.band_solve(*csc, B.T, perm=O.grid_transpose_perm(nx, ny))
        assert rel(X, Xr.T) <= 1e-10


def test_umfpack_abi_config_201x101():
    """201x101 grid, h=10 m, 20-cell PML, inclusion model, 5 Hz, 10 sources / 100
    receivers on row 5: the reference's per-shot loop through the drop-in ABI."""
    w = models.config("ac_201x101_inclusion")
    nx, ny, n = w.nx, w.ny, w.nx * w.ny
    say("umfpack ABI ac_201x101_inclusion")
    tri, csc = oracle_system(G.AC, nx, ny, w.npml, w.dh, w.freqs[0], w.m1, w.m2)
    irow, icol, ar, ai = tri
    nnz = len(ar)
    L = G.lib()
    Ap = np.zeros(n + 1, np.int32)
    Ai = np.zeros(nnz, np.int32)
    Ax = np.zeros(nnz)
    Az = np.zeros(nnz)
    assert L.umfpack_zi_triplet_to_col(n, n, nnz, irow.ctypes.data, icol.ctypes.data, ar.ctypes.data,
                                       ai.ctypes.data, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data,
                                       Az.ctypes.data, None) == 0
    Ax[:] = ar                                 # the host's overwrite (src/AC/forward_shot_AC.c:43-48)
    Az[:] = ai
    assert np.array_equal(Ap, csc[0]) and np.array_equal(Ai, csc[1])
    sym, num = C.c_void_p(), C.c_void_p()
    info = np.zeros(90)
    assert L.umfpack_zi_symbolic(n, n, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, Az.ctypes.data,
                                 C.byref(sym), None, None) == 0
    assert L.umfpack_zi_numeric(Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, Az.ctypes.data, sym,
                                C.byref(num), None, None) == 0
    L.umfpack_zi_free_symbolic(C.byref(sym))
    amp = O.source_amp(O.AC, w.freqs[0], 0.0, w.dh)
    B = np.zeros((n, len(w.src)), complex)
    for s, k in enumerate(w.src):
        B[k, s] = amp
    Xr, _ = O.band_solve(*csc, B, perm=O.grid_transpose_perm(nx, ny))
    xr, xi = np.zeros(n + 1), np.zeros(n + 1)      # the host allocates more than n (src/AC/forward_shot_AC.c:34-35)
    for s in range(len(w.src)):
        br, bi = np.ascontiguousarray(B[:, s].real), np.ascontiguousarray(B[:, s].imag)
        st = L.umfpack_zi_solve(0, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, Az.ctypes.data, xr.ctypes.data,
                                xi.ctypes.data, br.ctypes.data, bi.ctypes.data, num, None, info.ctypes.data)
        assert st == 0
        x = xr[:n] + 1j * xi[:n]
        assert rel(x, Xr[:, s]) <= 1e-10
        assert rel(x[w.rec], Xr[w.rec, s]) <= 1e-10
        assert info[80] <= 2                   # refinement steps taken
    L.umfpack_zi_free_numeric(C.byref(num))
    assert num.value is None


def test_forward_shots_fp32_refined_config_201x101():
    w = models.config("ac_201x101_inclusion")
    say("forward_shots ac_201x101_inclusion fp32+IR")
    pl = fdfd.Plan(w.nx, w.ny)
    out = fdfd.forward_shots(pl, w.physics, w.dh, w.freqs, dev(w.m1), dev(w.m2), w.src, w.rec, npml=w.npml,
                             precision=G.FP32, tol=1e-10, keep_fields=True)
    _, csc = oracle_system(G.AC, w.nx, w.ny, w.npml, w.dh, w.freqs[0], w.m1, w.m2)
    amp = O.source_amp(O.AC, w.freqs[0], 0.0, w.dh)
    B = np.zeros((w.nx * w.ny, len(w.src)), complex)
    for s, k in enumerate(w.src):
        B[k, s] = amp
    Xr, _ = O.band_solve(*csc, B, perm=O.grid_transpose_perm(w.nx, w.ny))
    X = out["fields"][0].cpu().numpy().T
    say("  iterations", out["iters"], "relres", out["relres"])
    assert out["relres"][0] <= 1e-10
    err = np.linalg.norm(X - Xr, axis=0) / np.linalg.norm(Xr, axis=0)
    assert np.all(err <= 1e-8), err
    d = out["data"][0].cpu().numpy()
    np.testing.assert_allclose(d, Xr[w.rec, :].T.astype(np.complex64), rtol=1e-5, atol=1e-6 * np.abs(d).max())


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
@pytest.mark.parametrize("sysno", [0, 2])
def test_batched_solves_match_oracle(precision, sysno):
    """Several frequencies in one launch set (and refinement that drops converged
    frequencies) against per-frequency oracle solves."""
    nx, ny, npml, dh = 41, 30, 6, 10.0
    m1, m2 = smooth_model(nx, ny, seed=4)
    freqs = [2.5, 7.0, 13.0]
    say("batched solves", precision, sysno)
    pl = fdfd.Plan(nx, ny, small_max=16)
    F = pl.factor(pl.assemble(G.AC, dh, freqs, dev(m1), dev(m2), npml=npml), precision)
    rng = np.random.default_rng(11)
    B = rng.standard_normal((3, 66, nx * ny)) + 1j * rng.standard_normal((3, 66, nx * ny))
    order = [2, 0, 1]
    if precision == G.FP64:
        X = F.solve_batch(order, dev(B), sys=sysno).cpu().numpy()
        tol = 1e-10
    else:
        X, its, rr = F.solve_refine_batch(order, dev(B), sys=sysno, tol=1e-12, max_iter=20)
        X = X.cpu().numpy()
        say("  iterations", its, rr)
        assert max(rr) <= 1e-12
        tol = 1e-8
    for z, k in enumerate(order):
        _, csc = oracle_system(G.AC, nx, ny, npml, dh, freqs[k], m1, m2)
        Xr, _ = O.band_solve(*csc, B[z].T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        assert rel(X[z], Xr.T) <= tol, (z, k)


@pytest.mark.parametrize("precision", [G.FP32, G.FP64])
@pytest.mark.parametrize("sysno", [0, 1, 2])
def test_refinement_layouts_agree(precision, sysno):
    """The panel refinement in elimination order (the FP64 residual updated by
    r -= op(M) d per step, the corrections summed once at the end) against the
    refinement in the caller's layout (r = b - op(M) x from the FP64 iterate, a
    transpose before and after every correction): the two residuals differ by
    FP64 rounding only, so with the same number of steps the iterates agree to
    1e-13; both within 1e-8 of the oracle.  66 right-hand sides leave padding
    columns in both internal row strides."""
    nx, ny, npml, dh = 41, 30, 6, 10.0
    m1, m2 = smooth_model(nx, ny, seed=4)
    freqs = [2.5, 7.0, 13.0]
    say("refinement layouts", precision, sysno)
    pl = fdfd.Plan(nx, ny, small_max=16)
    F = pl.factor(pl.assemble(G.AC, dh, freqs, dev(m1), dev(m2), npml=npml), precision)
    rng = np.random.default_rng(17 + sysno)
    B = dev(rng.standard_normal((3, 66, nx * ny)) + 1j * rng.standard_normal((3, 66, nx * ny)))
    order = [1, 2, 0]
    out = {}
    try:
        for layout in (0, 1):
            fdfd.solve_tuning(refine_layout=layout)
            X, its, rr = F.solve_refine_batch(order, B, sys=sysno, tol=1e-12, max_iter=30)
            out[layout] = (X.cpu().numpy(), list(its), list(rr))
            say("  layout", layout, "iterations", its, rr)
    finally:
        fdfd.solve_tuning()
    (X0, it0, rr0), (X1, it1, rr1) = out[0], out[1]
    assert max(rr0) <= 1e-12 and max(rr1) <= 1e-12
    if it0 == it1:
        assert rel(X0, X1) <= 1e-13
    Bh = B.cpu().numpy()
    for z, k in enumerate(order):
        _, csc = oracle_system(G.AC, nx, ny, npml, dh, freqs[k], m1, m2)
        Xr, _ = O.band_solve(*csc, Bh[z].T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        assert rel(X0[z], Xr.T) <= 1e-8 and rel(X1[z], Xr.T) <= 1e-8, (z, k)


def random_grid_csc(nx, ny, seed):
    """Random complex values on the 9-point pattern (no diagonal dominance): the
    in-front partial pivoting swaps rows, which the assembled Helmholtz matrices
    of the small tests rarely do."""
    pl = fdfd.Plan(nx, ny)
    cp, ri = pl.pattern()
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(len(ri)) + 1j * rng.standard_normal(len(ri))
    return cp, ri, np.ascontiguousarray(v.real), np.ascontiguousarray(v.imag)


@pytest.mark.parametrize("sysno", [0, 1, 2])
@pytest.mark.parametrize("nrhs", [1, 70])
def test_pivoting_solves_match_oracle(sysno, nrhs):
    nx, ny = 37, 26
    say("pivoting solves", sysno, nrhs)
    mats = [random_grid_csc(nx, ny, seed) for seed in (21, 22)]
    pl = fdfd.Plan(nx, ny, small_max=16)
    vals = dev(np.stack([c[2] + 1j * c[3] for c in mats]))
    F = pl.factor(vals, G.FP64)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((2, nrhs, nx * ny)) + 1j * rng.standard_normal((2, nrhs, nx * ny))
    Xb = F.solve_batch([1, 0], dev(B), sys=sysno).cpu().numpy()
    for z, k in enumerate([1, 0]):
        Xr, _ = O.band_solve(*mats[k], B[z].T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        assert rel(Xb[z], Xr.T) <= 1e-8, (z, k)
        Xs = F.solve(k, dev(B[z]), sys=sysno).cpu().numpy()
        assert rel(Xs, Xr.T) <= 1e-8, (z, k)


@pytest.mark.parametrize("matrix", ["helmholtz", "random"])
@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_persistent_subtrees_equal_level_path(precision, matrix):
    """The small subtrees factored by the one persistent launch (a task is a
    whole subtree of one frequency) against the level-by-level
    kernels: the same arithmetic in the same order, so the solutions agree
    bit for bit, for assembled Helmholtz matrices and for random values whose
    pivoting swaps rows, three frequencies in one batch; and both against the
    banded oracle."""
    nx, ny = 61, 41
    say("persistent subtrees", precision, matrix)
    if matrix == "helmholtz":
        m1, m2 = smooth_model(nx, ny, seed=7)
        freqs = [3.0, 8.0, 13.5]
        mats = [oracle_system(G.AC, nx, ny, 6, 10.0, f, m1, m2)[1] for f in freqs]
    else:
        mats = [random_grid_csc(nx, ny, seed) for seed in (41, 42, 43)]
    pl = fdfd.Plan(nx, ny)
    vals = dev(np.stack([c[2] + 1j * c[3] for c in mats]))
    rng = np.random.default_rng(4)
    B = rng.standard_normal((3, 2, nx * ny)) + 1j * rng.standard_normal((3, 2, nx * ny))
    out = {}
    for persistent in (1, 0):
        fdfd.factor_tuning(0, persistent)
        try:
            F = pl.factor(vals, precision)
            assert np.all(F.singular() == 0)
            if precision == G.FP64:
                out[persistent] = F.solve_batch([0, 1, 2], dev(B)).cpu().numpy()
            else:
                out[persistent] = F.solve_refine_batch([0, 1, 2], dev(B), tol=1e-12, max_iter=30)[0].cpu().numpy()
        finally:
            fdfd.factor_tuning(0, 1)
    assert np.array_equal(out[0], out[1])
    for z in range(3):
        Xr, _ = O.band_solve(*mats[z], B[z].T, sys=0, perm=O.grid_transpose_perm(nx, ny))
        assert rel(out[1][z], Xr.T) <= (1e-10 if matrix == "helmholtz" and precision == G.FP64 else 1e-8)


@pytest.mark.parametrize("matrix", ["helmholtz", "random"])
@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_two_panel_schedule_matches_oracle(precision, matrix):
    """The blocked path's two schedules -- two 32-column panels per step with
    one depth-64 trailing update (every level), and one panel per step with a
    depth-32 update (no level) -- each against the banded oracle for M and
    M^T, three frequencies in one batch, on a grid whose separators have more
    than 64 pivots (several 64-column blocks, a last block of one panel), on the
    assembled Helmholtz matrices and on random values whose partial pivoting
    swaps rows (so the second panel's swaps reach the deferred update)."""
    nx, ny = 97, 81
    say("two-panel schedule", precision, matrix)
    if matrix == "helmholtz":
        m1, m2 = smooth_model(nx, ny, seed=12)
        mats = [oracle_system(G.AC, nx, ny, 8, 10.0, f, m1, m2)[1] for f in (3.0, 9.0, 15.0)]
    else:
        mats = [random_grid_csc(nx, ny, seed) for seed in (51, 52, 53)]
    pl = fdfd.Plan(nx, ny, small_max=16)
    assert pl.info.max_pivots > 64
    vals = dev(np.stack([c[2] + 1j * c[3] for c in mats]))
    rng = np.random.default_rng(21)
    B = rng.standard_normal((3, 2, nx * ny)) + 1j * rng.standard_normal((3, 2, nx * ny))
    out = {}
    for deep in (0, fdfd.NO_DEEP):
        fdfd.factor_tuning(0, 1, deep)
        try:
            F = pl.factor(vals, precision)
        finally:
            fdfd.factor_tuning()
        assert np.all(F.singular() == 0)
        for sysno in (0, 2):
            if precision == G.FP64:
                X = F.solve_batch([0, 1, 2], dev(B), sys=sysno).cpu().numpy()
            else:
                X = F.solve_refine_batch([0, 1, 2], dev(B), sys=sysno, tol=1e-12, max_iter=30)[0].cpu().numpy()
            out[deep, sysno] = X
            for z in range(3):
                Xr, _ = O.band_solve(*mats[z], B[z].T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
                err = rel(X[z], Xr.T)
                say("  deep_min_m", deep, "sys", sysno, "freq", z, "rel err", err)
                assert err <= (1e-10 if matrix == "helmholtz" and precision == G.FP64 else 1e-8)
    # the two schedules give the same factors to rounding
    for sysno in (0, 2):
        assert rel(out[0, sysno], out[fdfd.NO_DEEP, sysno]) <= (1e-9 if precision == G.FP64 else 1e-6)


@pytest.mark.parametrize("matrix", ["helmholtz", "random"])
@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_global_memory_panel_path_matches_oracle(precision, matrix):
    """The blocked path's global-memory panel kernel, which factors panels too
    large for LDS (separators above 318 rows in FP64, 637 in FP32): forced on
    every blocked panel of a small grid by a one-byte LDS cap, against the
    banded oracle for M and M^T -- on the assembled Helmholtz matrix and on
    random values whose partial pivoting swaps rows."""
    nx, ny = 45, 33
    say("global-memory panel path", precision, matrix)
    if matrix == "helmholtz":
        _, _, csc = system(nx, ny)
    else:
        csc = random_grid_csc(nx, ny, 31)
    pl = fdfd.Plan(nx, ny, small_max=16)
    vals = dev((csc[2] + 1j * csc[3])[None, :])
    rng = np.random.default_rng(9)
    B = rng.standard_normal((3, nx * ny)) + 1j * rng.standard_normal((3, nx * ny))
    fdfd.factor_tuning(1)
    try:
        F = pl.factor(vals, precision)
    finally:
        fdfd.factor_tuning(0, 1)
    F0 = pl.factor(vals, precision)
    for sysno in (0, 2):
        Xr, _ = O.band_solve(*csc, B.T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        if precision == G.FP64:
            X = F.solve(0, dev(B), sys=sysno).cpu().numpy()
            X0 = F0.solve(0, dev(B), sys=sysno).cpu().numpy()
            bar = 1e-10 if matrix == "helmholtz" else 1e-8
        else:
            X, it, rr = F.solve_refine(0, dev(B), sys=sysno, tol=1e-12, max_iter=30)
            X0, _, _ = F0.solve_refine(0, dev(B), sys=sysno, tol=1e-12, max_iter=30)
            X, X0 = X.cpu().numpy(), X0.cpu().numpy()
            say("  sys", sysno, "iterations", it, "relres", rr)
            bar = 1e-8
        say("  sys", sysno, "rel err", rel(X, Xr.T), "vs LDS panels", rel(X, X0))
        assert rel(X, Xr.T) <= bar
        # the same factorisation as the LDS panel kernels, to rounding
        assert rel(X, X0) <= (1e-9 if precision == G.FP64 else 1e-6)


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
@pytest.mark.parametrize("path", [(64, -1), (16, -1), (32, 0)])
@pytest.mark.parametrize("sysno", [0, 1, 2])
def test_solve_paths_match_oracle(precision, path, sysno):
    """90x64 grid: separators of up to 64 unknowns make multi-block fronts
    above the single-block levels.  path = (widest fused tile, largest
    multi-block front order on the fused kernels; 0 sends those levels to the
    per-block kernels)."""
    nx, ny, npml, dh = 90, 64, 8, 10.0
    m1, m2 = smooth_model(nx, ny, seed=7)
    freqs = [4.0, 11.0]
    say("solve paths", precision, path, sysno)
    mats = [oracle_system(G.AC, nx, ny, npml, dh, f, m1, m2)[1] for f in freqs]
    mats.append(random_grid_csc(nx, ny, 31 + sysno))
    vals = dev(np.stack([c[2] + 1j * c[3] for c in mats]))
    pl = fdfd.Plan(nx, ny)
    assert pl.info.max_pivots > 32
    rng = np.random.default_rng(5 + sysno)
    B = rng.standard_normal((3, 40, nx * ny)) + 1j * rng.standard_normal((3, 40, nx * ny))
    fdfd.solve_tuning(*path)
    try:
        F = pl.factor(vals, precision)
        X1 = None
        if precision == G.FP64:
            X = F.solve_batch([0, 1, 2], dev(B), sys=sysno).cpu().numpy()
            tol = 1e-10
        else:
            # the unrefined FP32 solve is checked on its own as well: refinement
            # would hide a wrong substitution behind extra iterations
            X1 = F.solve_batch([0, 1, 2], dev(B), sys=sysno).cpu().numpy()
            X, its, rr = F.solve_refine_batch([0, 1, 2], dev(B), sys=sysno, tol=1e-12, max_iter=30)
            X = X.cpu().numpy()
            say("  iterations", its, rr)
            tol = 1e-8
    finally:
        fdfd.solve_tuning()
    for z, csc in enumerate(mats):
        Xr, info = O.band_solve(*csc, B[z].T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        assert info == 0
        assert rel(X[z], Xr.T) <= (tol if z < 2 else 100 * tol), z
        if X1 is not None and z < 2:
            e1 = rel(X1[z], Xr.T)
            say("  unrefined FP32 error", z, e1)
            assert e1 <= 1e-3, (z, e1)


@pytest.mark.parametrize("sysno", [0, 2])
def test_refinement_folds_corrections(sysno):
    """More corrections than the elimination-order refinement keeps apart (8):
    with tol 0 every frequency runs max_iter = 17 steps, so the corrections are
    folded into the FP64 sum twice (before steps 7 and 15) and the final pass
    adds that sum to the two corrections left.  Against the caller's-layout
    refinement after the same 17 steps (which carries x itself): 1e-12; both
    within 1e-8 of the oracle."""
    nx, ny, npml, dh = 41, 30, 6, 10.0
    m1, m2 = smooth_model(nx, ny, seed=5)
    freqs = [3.0, 9.0]
    say("refinement folds", sysno)
    pl = fdfd.Plan(nx, ny, small_max=16)
    F = pl.factor(pl.assemble(G.AC, dh, freqs, dev(m1), dev(m2), npml=npml), G.FP32)
    rng = np.random.default_rng(31 + sysno)
    B = dev(rng.standard_normal((2, 40, nx * ny)) + 1j * rng.standard_normal((2, 40, nx * ny)))
    out = {}
    try:
        for layout in (0, 1):
            fdfd.solve_tuning(refine_layout=layout)
            X, its, rr = F.solve_refine_batch([1, 0], B, sys=sysno, tol=0.0, max_iter=17)
            out[layout] = X.cpu().numpy()
            say("  layout", layout, "iterations", list(its), list(rr))
            assert list(its) == [17, 17] and max(rr) <= 1e-12
    finally:
        fdfd.solve_tuning()
    assert rel(out[0], out[1]) <= 1e-12
    Bh = B.cpu().numpy()
    for z, k in enumerate([1, 0]):
        _, csc = oracle_system(G.AC, nx, ny, npml, dh, freqs[k], m1, m2)
        Xr, _ = O.band_solve(*csc, Bh[z].T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        assert rel(out[0][z], Xr.T) <= 1e-8 and rel(out[1][z], Xr.T) <= 1e-8, (z, k)


@pytest.mark.parametrize("sysno", [0, 2])
def test_refinement_packs_unconverged_columns(sysno):
    """Correction solves on the columns still above the tolerance only: 30 of
    100 right-hand sides are nonzero (the zero ones have converged from the
    start), so every correction solve of the elimination-order refinement runs
    on a packed 32-column panel and the other columns get zero corrections.
    The zero columns stay exactly zero; against the caller's-layout
    refinement (which solves every column) the iterates agree to 1e-12 when
    the step counts agree; both within 1e-8 of the oracle."""
    nx, ny, npml, dh = 41, 30, 6, 10.0
    m1, m2 = smooth_model(nx, ny, seed=6)
    freqs = [4.0, 11.0]
    say("refinement packs columns", sysno)
    pl = fdfd.Plan(nx, ny, small_max=16)
    F = pl.factor(pl.assemble(G.AC, dh, freqs, dev(m1), dev(m2), npml=npml), G.FP32)
    rng = np.random.default_rng(41 + sysno)
    live = np.sort(rng.choice(100, 30, replace=False))
    Bh = np.zeros((2, 100, nx * ny), complex)
    Bh[:, live] = rng.standard_normal((2, 30, nx * ny)) + 1j * rng.standard_normal((2, 30, nx * ny))
    B = dev(Bh)
    out = {}
    try:
        for layout in (0, 1):
            fdfd.solve_tuning(refine_layout=layout)
            X, its, rr = F.solve_refine_batch([1, 0], B, sys=sysno, tol=1e-12, max_iter=30)
            out[layout] = (X.cpu().numpy(), list(its))
            say("  layout", layout, "iterations", its, rr)
            assert max(rr) <= 1e-12
    finally:
        fdfd.solve_tuning()
    (X0, it0), (X1, it1) = out[0], out[1]
    dead = np.setdiff1d(np.arange(100), live)
    assert np.all(X0[:, dead] == 0)
    if it0 == it1:
        assert rel(X0, X1) <= 1e-12
    for z, k in enumerate([1, 0]):
        _, csc = oracle_system(G.AC, nx, ny, npml, dh, freqs[k], m1, m2)
        Xr, _ = O.band_solve(*csc, Bh[z][live].T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
        assert rel(X0[z][live], Xr.T) <= 1e-8 and rel(X1[z][live], Xr.T) <= 1e-8, (z, k)


def test_refinement_wide_panel():
    """The fused elimination-order refinement at 1000 right-hand sides (near its
    1024 limit, where the per-wave column sums need more than the default 64 KB
    of LDS) on an odd grid height (the two-row tiles end half empty): within
    1e-8 of the oracle."""
    nx, ny, npml, dh = 37, 23, 5, 10.0
    m1, m2 = smooth_model(nx, ny, seed=9)
    freqs = [6.0]
    say("wide refinement panel")
    pl = fdfd.Plan(nx, ny)
    F = pl.factor(pl.assemble(G.AC, dh, freqs, dev(m1), dev(m2), npml=npml), G.FP32)
    rng = np.random.default_rng(23)
    B = rng.standard_normal((1, 1000, nx * ny)) + 1j * rng.standard_normal((1, 1000, nx * ny))
    X, its, rr = F.solve_refine_batch([0], dev(B), tol=1e-12, max_iter=30)
    say("  iterations", its, rr)
    assert max(rr) <= 1e-12
    _, csc = oracle_system(G.AC, nx, ny, npml, dh, freqs[0], m1, m2)
    Xr, _ = O.band_solve(*csc, B[0].T, perm=O.grid_transpose_perm(nx, ny))
    assert rel(X.cpu().numpy()[0], Xr.T) <= 1e-8


def test_refactor_equals_fresh_factorisation():
    nx, ny, npml, dh = 61, 41, 6, 10.0
    m1, m2 = smooth_model(nx, ny, seed=9)
    say("refactor")
    pl = fdfd.Plan(nx, ny)
    va = pl.assemble(G.AC, dh, [3.0, 8.0], dev(m1), dev(m2), npml=npml)
    vb = pl.assemble(G.AC, dh, [5.0, 12.0], dev(m1), dev(m2), npml=npml)
    rng = np.random.default_rng(2)
    B = dev(rng.standard_normal((2, 8, nx * ny)) + 1j * rng.standard_normal((2, 8, nx * ny)))
    for precision in (G.FP64, G.FP32):
        F = pl.factor(va, precision).refactor(vb)
        Xa = F.solve_batch([0, 1], B).cpu().numpy()
        Xb = pl.factor(vb, precision).solve_batch([0, 1], B).cpu().numpy()
        assert rel(Xa, Xb) <= 1e-13


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_refactor_graph_replays_match_eager(precision):
    """The refactorisation captured as one graph and replayed: every replay, with
    eager work and synchronisations between replays and the values buffer
    refilled in between, equals the eager refactorisation of the same values
    bit for bit (same kernels, same order), and the oracle to 1e-10 (FP64)."""
    nx, ny, npml, dh = 61, 41, 6, 10.0
    m1, m2 = smooth_model(nx, ny, seed=9)
    say("refactor graph", precision)
    pl = fdfd.Plan(nx, ny)
    fa, fb = [3.0, 8.0], [5.0, 12.0]
    va = pl.assemble(G.AC, dh, fa, dev(m1), dev(m2), npml=npml)
    vb = pl.assemble(G.AC, dh, fb, dev(m1), dev(m2), npml=npml)
    rng = np.random.default_rng(4)
    B = dev(rng.standard_normal((2, 8, nx * ny)) + 1j * rng.standard_normal((2, 8, nx * ny)))
    ref = {}
    for name, v in (("a", va), ("b", vb)):
        ref[name] = pl.factor(v, precision).solve_batch([0, 1], B).cpu().numpy()
    buf = torch.empty_like(va)
    F = pl.factor(va, precision)
    for k, name in enumerate(["b", "a", "b", "b", "a"]):
        buf.copy_(vb if name == "b" else va)
        F.refactor_graph(buf)
        X = F.solve_batch([0, 1], B).cpu().numpy()          # eager solves between the replays
        torch.cuda.synchronize()
        assert np.array_equal(X, ref[name]), (k, name)
    if precision == G.FP64:
        for z, f in enumerate(fa):
            _, csc = oracle_system(G.AC, nx, ny, npml, dh, f, m1, m2)
            Xr, _ = O.band_solve(*csc, B[z].cpu().numpy().T, perm=O.grid_transpose_perm(nx, ny))
            assert rel(X[z], Xr.T) <= 1e-10


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_single_vector_wide_levels_match_panel_path(precision):
    """A grid whose top separators exceed 256 rows, so the single-vector
    backward products of its top levels take the 1024-thread workgroups and
    the lower multi-block levels the 16- and 32-row groups: single right-hand
    sides (the one-wave and split-product kernels) against the same columns
    solved as one panel of five (the multi-right-hand-side kernels, checked
    against the oracle elsewhere) on the same factors, for M, M^H and M^T,
    and the FP64 backward error of the single-vector solutions.  FP64: 1e-10
    apart, backward error 1e-11 (max norms; measured 4e-13 to 8e-13); FP32
    factors with FP64 refinement to 1e-12: 1e-8 apart, backward error 1e-10."""
    nx, ny = 601, 301
    say("single-vector wide levels", precision)
    m1, m2 = smooth_model(nx, ny, seed=21)
    pl = fdfd.Plan(nx, ny)
    assert pl.info.max_pivots > 256
    vals = pl.assemble(G.AC, 10.0, [9.0], dev(m1), dev(m2), npml=10)
    F = pl.factor(vals, precision)
    rng = np.random.default_rng(8)
    B = dev(rng.standard_normal((5, pl.n)) + 1j * rng.standard_normal((5, pl.n)))
    for sysno in (0, 1, 2):
        if precision == G.FP64:
            Xp = F.solve(0, B, sys=sysno)
            Xv = [F.solve(0, B[c:c + 1], sys=sysno) for c in (0, 3)]
            bar, bbar = 1e-10, 1e-11
        else:
            Xp = F.solve_refine(0, B, sys=sysno, tol=1e-12, max_iter=30)[0]
            Xv = [F.solve_refine(0, B[c:c + 1], sys=sysno, tol=1e-12, max_iter=30)[0] for c in (0, 3)]
            bar, bbar = 1e-8, 1e-10
        for c, X in zip((0, 3), Xv):
            e = rel(X[0].cpu().numpy(), Xp[c].cpu().numpy())
            R = pl.residual(vals[0], B[c:c + 1], X, sys=sysno)
            be = float(R.abs().max() / B[c].abs().max())
            say("  sys", sysno, "column", c, "vs panel", e, "backward error", be)
            assert e <= bar
            assert be <= bbar


@pytest.mark.parametrize("matrix", ["helmholtz", "random"])
@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_single_vector_inverse_path_matches_oracle(precision, matrix):
    """One right-hand side at a time on a grid whose top levels hold a few
    multi-block fronts (pivot blocks of up to 81 rows): those levels run on
    the inverse pivot triangles (made once after the factorisation) and the
    split products, for M, M^H and M^T, against the banded oracle; the same
    factors solved again after a refactorisation with other values (the
    inverses are remade)."""
    nx, ny = 97, 81
    say("single-vector inverse path", precision, matrix)
    if matrix == "helmholtz":
        m1, m2 = smooth_model(nx, ny, seed=14)
        mats = [oracle_system(G.AC, nx, ny, 8, 10.0, f, m1, m2)[1] for f in (4.0, 13.0)]
    else:
        mats = [random_grid_csc(nx, ny, seed) for seed in (61, 62)]
    pl = fdfd.Plan(nx, ny, small_max=16)
    assert pl.info.max_pivots > 64
    rng = np.random.default_rng(5)
    B = rng.standard_normal((1, nx * ny)) + 1j * rng.standard_normal((1, nx * ny))
    F = None
    for z, csc in enumerate(mats):
        vals = dev((csc[2] + 1j * csc[3])[None, :])
        F = pl.factor(vals, precision) if F is None else F.refactor(vals)
        for sysno in (0, 1, 2):
            Xr, _ = O.band_solve(*csc, B.T, sys=sysno, perm=O.grid_transpose_perm(nx, ny))
            if precision == G.FP64:
                X = F.solve(0, dev(B), sys=sysno).cpu().numpy()
                bar = 1e-10 if matrix == "helmholtz" else 1e-8
            else:
                X = F.solve_refine(0, dev(B), sys=sysno, tol=1e-12, max_iter=30)[0].cpu().numpy()
                bar = 1e-8
            e = rel(X, Xr.T)
            say("  matrix", z, "sys", sysno, "rel err", e)
            assert e <= bar
