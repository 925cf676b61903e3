"""ORACLE — test infrastructure only.

ctypes front-end of ``liboracle.so`` (oracle/assemble.c, oracle/bandlu.c): the
CPU restatement of GERMAINE's assembly, shot kernels and an FP64 banded LU
checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module; the product package never does.

Conventions: grids are nx*ny including the PML; material arrays are float32
``[ny, nx]``; matrix values are returned in the reference's triplet order
(row k = (i-1)+(j-1)*nx, columns ascending), which is also the CSC order of
the matrix M = A^T that UMFPACK factors (src/AC/forward_shot_AC.c:41-48).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

AC, SH, TE = 1, 2, 3


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        _LIB = C.CDLL(path)
        f = np.ctypeslib.ndpointer
        L = _LIB
        L.orc_nonzero.restype = C.c_int
        L.orc_nonzero.argtypes = [C.c_int, C.c_int]
        L.orc_pml_profiles.restype = None
        L.orc_pml_profiles.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.c_int, f(np.float32), f(np.float32)]
        L.orc_omega2.restype = None
        L.orc_omega2.argtypes = [C.c_int, C.c_float, C.c_float, f(np.float32), f(np.float32)]
        for name in ("orc_assemble_ac", "orc_assemble_sh", "orc_assemble_te"):
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                           f(np.float32), f(np.float32), f(np.float32), f(np.float32),
                           f(np.int32), f(np.int32), f(np.float64), f(np.float64)]
        L.orc_smith_divf.restype = None
        L.orc_smith_divf.argtypes = [C.c_float] * 4 + [C.POINTER(C.c_float)] * 2
        L.orc_c_divf.restype = None
        L.orc_c_divf.argtypes = [C.c_float] * 4 + [C.POINTER(C.c_float)] * 2
        L.orc_source_index.restype = C.c_int
        L.orc_source_index.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int]
        L.orc_source_amp.restype = None
        L.orc_source_amp.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_residual.restype = C.c_float
        L.orc_residual.argtypes = [C.c_int] + [f(np.float32)] * 6
        L.orc_residual_ex.restype = C.c_float
        L.orc_residual_ex.argtypes = [C.c_int] * 4 + [f(np.float32)] * 6 + [f(np.float32)]
        L.orc_adjoint_rhs.restype = None
        L.orc_adjoint_rhs.argtypes = [C.c_int, C.c_int, f(np.int32), f(np.float32), f(np.float32),
                                      C.c_float, f(np.float64), f(np.float64)]
        L.orc_grad_ac_accumulate.restype = None
        L.orc_grad_ac_accumulate.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float,
                                             f(np.float64), f(np.float64), C.c_double, C.c_double,
                                             f(np.float64), f(np.float64), C.c_double, C.c_double,
                                             f(np.float32), f(np.float32), C.c_float, C.c_float, f(np.float32)]
        L.orc_wavelet_amp.restype = None
        L.orc_wavelet_amp.argtypes = [C.c_int, C.c_float, C.c_float, f(np.float32), C.c_int, C.c_int,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_taper_grad_shot.restype = None
        L.orc_taper_grad_shot.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                          f(np.int32), f(np.int32), f(np.float32)]
        L.orc_taper_grad_shot1.restype = None
        L.orc_taper_grad_shot1.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, f(np.float32),
                                           f(np.float32), C.c_int, f(np.int32), f(np.int32), f(np.float32)]
        L.orc_hess_shin_ac_accumulate.restype = None
        L.orc_hess_shin_ac_accumulate.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float,
                                                  f(np.float64), f(np.float64), C.c_double, C.c_double,
                                                  f(np.float32), f(np.float32), f(np.float32)]
        L.orc_apply_hess.restype = None
        L.orc_apply_hess.argtypes = [C.c_int, C.c_float, f(np.float32), f(np.float32)]
        L.orc_model_update_norm.restype = None
        L.orc_model_update_norm.argtypes = [C.c_int, C.c_float, f(np.float32), f(np.float32), C.c_float, C.c_float,
                                            C.c_float, f(np.float32)]
        L.orc_precond.restype = None
        L.orc_precond.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                  C.c_int, f(np.float32), f(np.float32)]
        L.orc_tikhonov_te.restype = None
        L.orc_tikhonov_te.argtypes = [C.c_int, C.c_int] + [C.c_float] * 7 + [f(np.float32)] * 5
        L.orc_taper_file.restype = None
        L.orc_taper_file.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, f(np.float32), f(np.float32)]
        L.orc_grad_frame.restype = None
        L.orc_grad_frame.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, f(np.float32)]
        L.orc_model_update.restype = None
        L.orc_model_update.argtypes = [C.c_int, C.c_float, f(np.float32), f(np.float32), C.c_float, C.c_float,
                                       f(np.float32)]
        L.orc_hess_gn_ac_accumulate.restype = None
        L.orc_hess_gn_ac_accumulate.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float,
                                                f(np.float64), f(np.float64), C.c_double, C.c_double,
                                                f(np.float64), f(np.float64), C.c_double, C.c_double,
                                                f(np.float32), f(np.float32), f(np.float32)]
        L.orc_hess_gn_te_accumulate.restype = None
        L.orc_hess_gn_te_accumulate.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float,
                                                f(np.float64), f(np.float64), C.c_double, C.c_double,
                                                f(np.float64), f(np.float64), C.c_double, C.c_double,
                                                f(np.float32), f(np.float32)]
        L.orc_hess_shin_te_accumulate.restype = None
        L.orc_hess_shin_te_accumulate.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float,
                                                  f(np.float64), f(np.float64), C.c_double, C.c_double,
                                                  f(np.float32), f(np.float32)]
        L.orc_grad_te_accumulate.restype = None
        L.orc_grad_te_accumulate.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float,
                                             f(np.float64), f(np.float64), C.c_double, C.c_double,
                                             f(np.float64), f(np.float64), C.c_double, C.c_double,
                                             C.c_float, C.c_float, f(np.float32), f(np.float32)]
        L.orc_band_solve.restype = C.c_int
        L.orc_band_solve.argtypes = [C.c_int, f(np.int32), f(np.int32), f(np.float64), f(np.float64),
                                     C.c_void_p, C.c_int, C.c_int, f(np.float64), f(np.float64),
                                     f(np.float64), f(np.float64)]
        L.orc_csc_matvec.restype = None
        L.orc_csc_matvec.argtypes = [C.c_int, f(np.int32), f(np.int32), f(np.float64), f(np.float64),
                                     C.c_int, f(np.float64), f(np.float64), f(np.float64), f(np.float64)]
    return _LIB


def nonzero(nx, ny):
    return lib().orc_nonzero(nx, ny)


def pml_profiles(nx, ny, npml, dh, a0_pml, freq, s=0.0, free_surf=0):
    """(px, py): float32 [4, nx+2] / [4, ny+2] rows = damp r, i, half r, half i."""
    px = np.zeros(4 * (nx + 2), np.float32)
    py = np.zeros(4 * (ny + 2), np.float32)
    lib().orc_pml_profiles(nx, ny, npml, dh, a0_pml, freq, s, free_surf, px, py)
    return px.reshape(4, nx + 2), py.reshape(4, ny + 2)


def omega2(physics, freq, s=0.0):
    o2 = np.zeros(2, np.float32)
    om = np.zeros(2, np.float32)
    lib().orc_omega2(physics, freq, s, o2, om)
    return o2, om


def assemble(physics, nx, ny, dh, freq, s, m1, m2, px, py):
    """Triplets (irow, icol, Ar, Ai) of A in the reference order.

    physics AC: m1=vp, m2=rho; SH: m1=vs, m2=rho; TE: m1=eps (absolute), m2=sigma.
    """
    nnz = nonzero(nx, ny)
    irow = np.zeros(nnz, np.int32)
    icol = np.zeros(nnz, np.int32)
    ar = np.zeros(nnz, np.float64)
    ai = np.zeros(nnz, np.float64)
    fn = {AC: lib().orc_assemble_ac, SH: lib().orc_assemble_sh, TE: lib().orc_assemble_te}[physics]
    cnt = fn(nx, ny, dh, freq, s,
             np.ascontiguousarray(m1, np.float32).ravel(), np.ascontiguousarray(m2, np.float32).ravel(),
             np.ascontiguousarray(px, np.float32).ravel(), np.ascontiguousarray(py, np.float32).ravel(),
             irow, icol, ar, ai)
    assert cnt == nnz
    return irow, icol, ar, ai


def csc_of_M(nx, ny, irow, icol, ar, ai):
    """CSC arrays (Ap, Ai, Ax, Az) of the matrix UMFPACK sees after the host's
    value overwrite (src/AC/forward_shot_AC.c:41-48): column l of M holds the
    triplets of row l of A, so M = A^T and Ax is the triplet value array."""
    n = nx * ny
    ap = np.zeros(n + 1, np.int32)
    np.add.at(ap, irow + 1, 1)
    ap = np.cumsum(ap).astype(np.int32)
    return ap, icol.astype(np.int32), ar.copy(), ai.copy()


def band_solve(ap, ai, ax, az, b, sys=0, perm=None):
    """Solve op(M) x = b with the banded-LU oracle; b: complex [n] or [n, nrhs]."""
    b = np.asarray(b)
    one = b.ndim == 1
    B = b.reshape(b.shape[0], -1)
    n, nrhs = B.shape
    bx = np.ascontiguousarray(B.real.T, np.float64).ravel()
    bz = np.ascontiguousarray(B.imag.T, np.float64).ravel()
    xx = np.zeros(n * nrhs)
    xz = np.zeros(n * nrhs)
    p = None
    if perm is not None:
        perm = np.ascontiguousarray(perm, np.int32)
        p = perm.ctypes.data
    info = lib().orc_band_solve(n, ap, ai, ax, az, p, sys, nrhs, bx, bz, xx, xz)
    if info < 0:
        raise MemoryError("band LU allocation failed")
    X = (xx + 1j * xz).reshape(nrhs, n).T
    return (X[:, 0] if one else X), info


def grid_transpose_perm(nx, ny):
    """Symmetric permutation that orders the grid column-wise (bandwidth ny+1)."""
    return (np.arange(nx)[:, None] + nx * np.arange(ny)[None, :]).ravel().astype(np.int32)


def matvec(ap, ai, ax, az, x, sys=0):
    n = len(ap) - 1
    x = np.asarray(x, np.complex128)
    yr = np.zeros(n)
    yi = np.zeros(n)
    lib().orc_csc_matvec(n, ap, ai, ax, az, sys, np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag), yr, yi)
    return yr + 1j * yi


def source_index(x, y, dh, npml, fsshift, nx):
    return lib().orc_source_index(x, y, dh, npml, fsshift, nx)


def source_amp(physics, freq, s, dh):
    re = C.c_double()
    im = C.c_double()
    lib().orc_source_amp(physics, freq, s, dh, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def residual(pcal, pobs):
    pcal = np.asarray(pcal, np.complex64)
    pobs = np.asarray(pobs, np.complex64)
    n = len(pcal)
    rr = np.zeros(n, np.float32)
    ri = np.zeros(n, np.float32)
    l2 = lib().orc_residual(n, np.ascontiguousarray(pcal.real), np.ascontiguousarray(pcal.imag),
                            np.ascontiguousarray(pobs.real), np.ascontiguousarray(pobs.imag), rr, ri)
    return rr + 1j * ri, l2


def residual_ex(pcal, pobs, misfit=1, stf_inv=0, rtm=0):
    """calc_res_AC for one shot (src/AC/calc_res_AC.c:10-95): returns
    (muted residuals, sum |res|^2, Wiener source-time-function estimate)."""
    pcal = np.asarray(pcal, np.complex64)
    pobs = np.asarray(pobs, np.complex64)
    n = len(pcal)
    rr = np.zeros(n, np.float32)
    ri = np.zeros(n, np.float32)
    stf = np.zeros(2, np.float32)
    l2 = lib().orc_residual_ex(n, misfit, stf_inv, rtm, np.ascontiguousarray(pcal.real),
                               np.ascontiguousarray(pcal.imag), np.ascontiguousarray(pobs.real),
                               np.ascontiguousarray(pobs.imag), rr, ri, stf)
    return rr + 1j * ri, l2, complex(stf[0], stf[1])


def adjoint_rhs(n, krec, res, dh):
    res = np.asarray(res, np.complex64)
    rr = np.zeros(n)
    ri = np.zeros(n)
    lib().orc_adjoint_rhs(n, len(krec), np.ascontiguousarray(krec, np.int32),
                          np.ascontiguousarray(res.real), np.ascontiguousarray(res.imag), dh, rr, ri)
    return rr + 1j * ri


def hess_shin_ac_accumulate(nx, ny, freq, s, u, u_last, vp, rho, hess):
    lib().orc_hess_shin_ac_accumulate(nx, ny, freq, s,
                                      np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag),
                                      u_last.real, u_last.imag,
                                      np.ascontiguousarray(vp, np.float32).ravel(),
                                      np.ascontiguousarray(rho, np.float32).ravel(), hess)
    return hess


def apply_hess(eps, hess, grad):
    """grad (float32, modified in place) *= 1 / (hess + eps max|hess|)"""
    lib().orc_apply_hess(len(grad), eps, np.ascontiguousarray(hess, np.float32), grad)
    return grad


def model_update(alpha, direction, m_old, lo, hi):
    """Trial model of a line-search step, float32 (src/calc_mat_change_wolfe.c:27-46)."""
    m_old = np.ascontiguousarray(m_old, np.float32).ravel()
    out = np.empty_like(m_old)
    lib().orc_model_update(len(m_old), alpha, np.ascontiguousarray(direction, np.float32).ravel(), m_old, lo, hi, out)
    return out


def model_update_norm(alpha, direction, m_old, norm, lo, hi):
    """TE trial model of a line-search step on normalised material, float32
    (src/TE/parabolicls_TE.c:30-31,64-69)."""
    m_old = np.ascontiguousarray(m_old, np.float32).ravel()
    out = np.empty_like(m_old)
    lib().orc_model_update_norm(len(m_old), alpha, np.ascontiguousarray(direction, np.float32).ravel(), m_old, norm,
                                lo, hi, out)
    return out


def precond(field, dh, vref, fc_low, fsx, fsy, spatfilter=1, nzero=0):
    """gauss_filt and taper_grad_hor on a float32 [ny, nx] field (src/precond.c:17-26)."""
    field = np.ascontiguousarray(field, np.float32)
    ny, nx = field.shape
    out = np.empty_like(field)
    lib().orc_precond(nx, ny, dh, vref, fc_low, fsx, fsy, spatfilter, nzero, field.ravel(), out.ravel())
    return out


def tikhonov_te(sig, eps, dh, beta, norm, lam, gs, ge):
    """Tikhonov cost (tikh_sigma, tikh_epsilon) of float32 [ny, nx] fields; gs, ge
    (float32, cells like the fields) receive the gradient terms in place."""
    sig = np.ascontiguousarray(sig, np.float32)
    ny, nx = sig.shape
    cost = np.zeros(2, np.float32)
    lib().orc_tikhonov_te(nx, ny, dh, beta[0], beta[1], norm[0], norm[1], lam[0], lam[1], sig.ravel(),
                          np.ascontiguousarray(eps, np.float32).ravel(), gs, ge, cost)
    return float(cost[0]), float(cost[1])


def taper_file(g, taper, npml, free_surf):
    """taper_grad_file on a float32 [ny, nx] field (a tapered copy is returned)."""
    taper = np.ascontiguousarray(taper, np.float32)
    ny0, nx0 = taper.shape
    out = np.ascontiguousarray(g, np.float32).copy()
    lib().orc_taper_file(nx0, ny0, npml, free_surf, taper.ravel(), out.ravel())
    return out


def grad_frame(g, nx0, ny0, npml, free_surf):
    """cp_grad_frame's PML damping on a float32 [ny, nx] field (a damped copy is returned)."""
    out = np.ascontiguousarray(g, np.float32).copy()
    lib().orc_grad_frame(nx0, ny0, npml, free_surf, out.ravel())
    return out


def hess_gn_ac_accumulate(nx, ny, freq, s, u, u_last, g, g_last, vp, rho, hess):
    lib().orc_hess_gn_ac_accumulate(nx, ny, freq, s,
                                    np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag),
                                    u_last.real, u_last.imag,
                                    np.ascontiguousarray(g.real), np.ascontiguousarray(g.imag),
                                    g_last.real, g_last.imag,
                                    np.ascontiguousarray(vp, np.float32).ravel(),
                                    np.ascontiguousarray(rho, np.float32).ravel(), hess)
    return hess


def hess_gn_te_accumulate(nx, ny, freq, s, u, u_last, g, g_last, hess_sigma, hess_eps):
    lib().orc_hess_gn_te_accumulate(nx, ny, freq, s,
                                    np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag),
                                    u_last.real, u_last.imag,
                                    np.ascontiguousarray(g.real), np.ascontiguousarray(g.imag),
                                    g_last.real, g_last.imag, hess_sigma, hess_eps)
    return hess_sigma, hess_eps


def hess_shin_te_accumulate(nx, ny, freq, s, u, u_last, hess_sigma, hess_eps):
    lib().orc_hess_shin_te_accumulate(nx, ny, freq, s,
                                      np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag),
                                      u_last.real, u_last.imag, hess_sigma, hess_eps)
    return hess_sigma, hess_eps


def grad_ac_accumulate(nx, ny, freq, s, u, u_last, lam, lam_last, vp, rho, grad, stf=1 + 0j):
    lib().orc_grad_ac_accumulate(nx, ny, freq, s,
                                 np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag),
                                 u_last.real, u_last.imag,
                                 np.ascontiguousarray(lam.real), np.ascontiguousarray(lam.imag),
                                 lam_last.real, lam_last.imag,
                                 np.ascontiguousarray(vp, np.float32).ravel(),
                                 np.ascontiguousarray(rho, np.float32).ravel(), stf.real, stf.imag, grad)
    return grad


def grad_te_accumulate(nx, ny, freq, s, u, u_last, lam, lam_last, gsig, geps, stf=1 + 0j):
    lib().orc_grad_te_accumulate(nx, ny, freq, s,
                                 np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag),
                                 u_last.real, u_last.imag,
                                 np.ascontiguousarray(lam.real), np.ascontiguousarray(lam.imag),
                                 lam_last.real, lam_last.imag, stf.real, stf.imag, gsig, geps)
    return gsig, geps


def wavelet_amp(physics, freq, dh, data, ndt):
    """Source amplitude of READ_WAVELET = 1: the DFT of the SU trace `data`
    (float32, ns samples; header interval ndt) at freq, over DH^2."""
    re, im = C.c_double(), C.c_double()
    d = np.ascontiguousarray(data, np.float32)
    lib().orc_wavelet_amp(1 if physics == TE else 0, freq, dh, d, len(d), int(ndt), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def taper_grad_shot(g, nx, ny, dh, srtradius, srcx, srcy, rec_i, rec_j):
    """taper_grad_shot on one shot's float32 gradient [ny*nx], in place."""
    lib().orc_taper_grad_shot(nx, ny, dh, srtradius, srcx, srcy, len(rec_i), np.ascontiguousarray(rec_i, np.int32),
                              np.ascontiguousarray(rec_j, np.int32), g)
    return g


def taper_grad_shot1(g, nx, ny, dh, filtsize, srcx, srcy, rec_i, rec_j):
    """taper_grad_shot1 on one shot's float32 gradient [ny*nx], in place; srcx/srcy:
    every source of the survey (metres)."""
    lib().orc_taper_grad_shot1(nx, ny, dh, filtsize, len(srcx), np.ascontiguousarray(srcx, np.float32),
                               np.ascontiguousarray(srcy, np.float32), len(rec_i),
                               np.ascontiguousarray(rec_i, np.int32), np.ascontiguousarray(rec_j, np.int32), g)
    return g
