"""Host-side workflows over libgfdfd.so, mirroring the reference's shot loops.

Device arrays are torch tensors on ``cuda:0`` (torch is used only to own
device memory and streams); every computation runs in the library's HIP
kernels.  Complex panels are column-major ``n x nshots``, i.e. torch tensors
of shape ``[nshots, n]``, dtype complex128.

* :func:`forward_shots` — ``forward_shot_AC`` (src/AC/forward_shot_AC.c:10-122)
  for a batch of frequencies and all shots at once: device assembly, batched
  factorisation, one multi-RHS panel solve per frequency, receiver sampling.
* :func:`grad_obj` — ``grad_obj_AC`` (src/AC/grad_obj_AC.c:10-194) /
  ``grad_obj_TE``: forward panel, residuals, adjoint panel on the same
  factors (M for AC, M^T for TE), gradient reduced over shots in FP64.
* :func:`objective`, :func:`model_update`, :func:`parabolic_line_search`,
  :func:`line_search` — ``obj_AC``, ``calc_mat_change_wolfe``, ``calc_opt_step``
  and ``parabolicls_AC``: the step search of every FWI iteration on one plan;
  :func:`model_update_norm`, :func:`line_search_te` — ``parabolicls_TE``.
* ``grad_obj(..., rtm=True)`` — ``RTM_FD_AC``: the gradient pass with the
  observed data as the residual, i.e. the migrated image; :func:`precond`
  smooths and tapers it (or an FWI gradient) as ``precond`` does.
"""
import ctypes as C

import numpy as np

from . import AC, FP32, FP64, SH, TE, FreqParams, PlanInfo, ShotTaperC, check, lib


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


# ---- host parameters (gcc arithmetic of the reference) -------------------------

def fd_weights(dh):
    out = np.zeros(6, np.float32)
    lib().gfdfd_fd_weights(dh, out.ctypes.data)
    return out


def pml_profiles(nx, ny, npml, dh, a0_pml, freq, s=0.0, free_surf=0):
    """src/pml_pro.c: float32 [4(nx+2) + 4(ny+2)] = x r,i,hr,hi then y r,i,hr,hi."""
    prof = np.zeros(4 * (nx + 2) + 4 * (ny + 2), np.float32)
    lib().gfdfd_pml_profiles(nx, ny, npml, dh, a0_pml, freq, s, free_surf, prof.ctypes.data)
    return prof


def omega(physics, freq, s=0.0):
    o2 = np.zeros(2, np.float32)
    om = np.zeros(2, np.float32)
    lib().gfdfd_omega(physics, freq, s, o2.ctypes.data, om.ctypes.data)
    return o2, om


def omega2_grad_ac(freq, s=0.0):
    o2 = np.zeros(2, np.float32)
    lib().gfdfd_omega2_grad_ac(freq, s, o2.ctypes.data)
    return o2


def source_amp(physics, freq, s, dh, wavelet=None):
    """Point-source amplitude of the right-hand side: the spike (src/AC/RHS_source_AC.c:37-53,
    src/TE/RHS_source_TE.c:22-56), or with wavelet = (samples, ndt) -- an SU trace as
    formats.read_su_wavelet returns it (READ_WAVELET = 1) -- its DFT at freq over DH^2
    (src/read_stf_dft.c:64-70)."""
    a = np.zeros(2, np.float64)
    if wavelet is None:
        lib().gfdfd_source_amp(physics, freq, s, dh, a.ctypes.data)
    else:
        w = np.ascontiguousarray(wavelet[0], np.float32)
        lib().gfdfd_wavelet_amp(physics, freq, dh, w.ctypes.data, len(w), int(wavelet[1]), a.ctypes.data)
    return complex(a[0], a[1])


def grid_index(x, y, dh, npml, fsshift, nx):
    """Row of a source/receiver at (x, y) metres: nxsrc + (nysrc-1)*NX with
    nxsrc = iround(x/DH)+NPML, nysrc = iround(y/DH)+FSSHIFT (src/AC/RHS_source_AC.c:23-32)."""
    xs = int(np.floor(np.float32(np.float32(x) / np.float32(dh)) + 0.5)) + npml
    ys = int(np.floor(np.float32(np.float32(y) / np.float32(dh)) + 0.5)) + fsshift
    return xs + (ys - 1) * nx


# ---- plan / factors ------------------------------------------------------------

def solve_tuning(rw_cap=64, fused_mmax=0, refine_layout=0):
    """Panel-solve kernel selection (process-wide): widest fused right-hand-side
    tile, the largest front order of a multi-block level that takes the
    fused per-front kernels (0, the default: all such levels take the
    per-block kernels; negative: as far as LDS allows), and the layout of the
    panel refinement (0: elimination order; 1: the caller's layout)."""
    check(lib().gfdfd_solve_tuning(rw_cap, fused_mmax, refine_layout), "gfdfd_solve_tuning")


NO_DEEP = 1 << 30   # deep_min_m that keeps every level on one panel per step


def factor_tuning(panel_lds_cap=0, persistent=1, deep_min_m=0):
    """Factorisation kernel selection (process-wide): panels needing more than
    panel_lds_cap bytes of LDS take the global-memory panel kernel (0: the
    hardware limit); persistent selects the one-launch factorisation of the
    small subtrees (else level by level); levels whose largest blocked front
    has order >= deep_min_m take the two-panel schedule (trailing update once
    per 64 pivots), the others one panel at a time (NO_DEEP: none)."""
    check(lib().gfdfd_factor_tuning(int(panel_lds_cap), int(persistent), int(deep_min_m)), "gfdfd_factor_tuning")


class Plan:
    """Nested-dissection plan of an nx x ny 9-point grid (cached by size)."""

    def __init__(self, nx, ny, leaf_area=32, small_max=64):
        self.nx, self.ny = nx, ny
        h = C.c_void_p()
        check(lib().gfdfd_plan_grid_opt(nx, ny, leaf_area, small_max, C.byref(h)), "gfdfd_plan_grid")
        self.h = h
        self.info = PlanInfo()
        check(lib().gfdfd_plan_query(self.h, C.byref(self.info)), "gfdfd_plan_query")
        self.n = int(self.info.n)
        self.nnz = int(self.info.nnz)

    def __del__(self):
        if getattr(self, "h", None) is not None and lib is not None:
            lib().gfdfd_plan_free(C.byref(self.h))
            self.h = None

    def pattern(self):
        cp = np.zeros(self.n + 1, np.int32)
        ri = np.zeros(self.nnz, np.int32)
        check(lib().gfdfd_plan_pattern(self.h, cp.ctypes.data, ri.ctypes.data), "gfdfd_plan_pattern")
        return cp, ri

    def perm(self):
        p = np.zeros(self.n, np.int32)
        check(lib().gfdfd_plan_perm(self.h, p.ctypes.data), "gfdfd_plan_perm")
        return p

    def fronts(self):
        nf = self.info.fronts
        arrs = [np.zeros(nf, np.int32) for _ in range(5)] + [np.zeros(nf, np.int64)]
        check(lib().gfdfd_plan_fronts(self.h, *[a.ctypes.data for a in arrs]), "gfdfd_plan_fronts")
        p, m, pos0, parent, level, idx_off = arrs
        idx = np.zeros(int(m.sum()), np.int32)
        check(lib().gfdfd_plan_index(self.h, idx.ctypes.data), "gfdfd_plan_index")
        return dict(p=p, m=m, pos0=pos0, parent=parent, level=level, idx_off=idx_off, idx=idx)

    def assemble(self, physics, dh, freqs, m1, m2, stream=None, s=0.0, npml=0, a0_pml=90.0, free_surf=0,
                 out=None):
        """Device assembly for a list of frequencies; m1/m2 float32 cuda [ny, nx].
        Returns complex128 cuda [nfreq, nnz] (values of M in CSC order), written
        into `out` when given."""
        import torch
        nf = len(freqs)
        profs, fp = [], (FreqParams * nf)()
        for k, f in enumerate(freqs):
            prof = pml_profiles(self.nx, self.ny, npml, dh, a0_pml, f, s, free_surf)
            profs.append(prof)
            o2, om = omega(physics, f, s)
            fp[k].omega2[0], fp[k].omega2[1] = float(o2[0]), float(o2[1])
            fp[k].omega[0], fp[k].omega[1] = float(om[0]), float(om[1])
            fp[k].prof = prof.ctypes.data_as(C.POINTER(C.c_float))
        vals = torch.empty((nf, self.nnz), dtype=torch.complex128, device="cuda") if out is None else out
        assert vals.shape == (nf, self.nnz) and vals.dtype == torch.complex128 and vals.is_contiguous()
        m1 = m1.contiguous()
        m2 = m2.contiguous()
        check(lib().gfdfd_assemble(self.h, physics, dh, nf, fp, _ptr(m1), _ptr(m2), _ptr(vals), _stream(stream)),
              "gfdfd_assemble")
        return vals

    def factor(self, vals, precision=FP64, stream=None):
        return Factors(self, vals, precision, stream)

    def residual(self, vals_f, B, X, sys=0, stream=None):
        import torch
        R = torch.empty_like(B)
        nrhs = B.shape[0] if B.dim() == 2 else 1
        check(lib().gfdfd_residual(self.h, _ptr(vals_f), sys, nrhs, _ptr(B), _ptr(X), _ptr(R), _stream(stream)),
              "gfdfd_residual")
        return R


class Factors:
    """Numeric factors of nfreq matrices (complex128 cuda [nfreq, nnz] values)."""

    def __init__(self, plan, vals, precision=FP64, stream=None):
        self.plan = plan
        self.nfreq = vals.shape[0]
        self.precision = precision
        vals = vals.contiguous()
        h = C.c_void_p()
        check(lib().gfdfd_factor(plan.h, self.nfreq, _ptr(vals), precision, _stream(stream), C.byref(h)),
              "gfdfd_factor")
        self.h = h

    def refactor(self, vals, stream=None):
        """New values (same pattern and frequency count) into these factors' storage."""
        vals = vals.contiguous()
        assert vals.shape[0] == self.nfreq
        check(lib().gfdfd_refactor(self.h, _ptr(vals), _stream(stream)), "gfdfd_refactor")
        return self

    def refactor_graph(self, vals, stream=None):
        """refactor as one hipGraph launch (captured on the first call with this
        values buffer, replayed afterwards: refill the same buffer between calls)."""
        assert vals.is_contiguous() and vals.shape[0] == self.nfreq
        check(lib().gfdfd_refactor_graph(self.h, _ptr(vals), _stream(stream)), "gfdfd_refactor_graph")
        return self

    def __del__(self):
        if getattr(self, "h", None) is not None and lib is not None:
            lib().gfdfd_factors_free(C.byref(self.h))
            self.h = None

    def singular(self):
        f = np.zeros(self.nfreq, np.int32)
        check(lib().gfdfd_factors_singular(self.h, f.ctypes.data), "gfdfd_factors_singular")
        return f

    def solve(self, ifreq, B, sys=0, out=None, stream=None):
        import torch
        X = torch.empty_like(B) if out is None else out
        nrhs = B.shape[0] if B.dim() == 2 else 1
        check(lib().gfdfd_solve(self.h, ifreq, sys, nrhs, _ptr(B), _ptr(X), _stream(stream)), "gfdfd_solve")
        return X

    def solve_refine(self, ifreq, B, sys=0, tol=1e-10, max_iter=10, out=None, stream=None):
        import torch
        X = torch.empty_like(B) if out is None else out
        nrhs = B.shape[0] if B.dim() == 2 else 1
        it = C.c_int(0)
        rr = C.c_double(0)
        check(lib().gfdfd_solve_refine(self.h, ifreq, sys, nrhs, _ptr(B), _ptr(X), tol, max_iter, C.byref(it),
                                       C.byref(rr), _stream(stream)), "gfdfd_solve_refine")
        return X, it.value, rr.value

    def solve_batch(self, freqs, B, sys=0, out=None, stream=None):
        """B: complex128 cuda [len(freqs), nrhs, n]; one launch set for all frequencies."""
        import torch
        X = torch.empty_like(B) if out is None else out
        fr = np.ascontiguousarray(freqs, np.int32)
        check(lib().gfdfd_solve_batch(self.h, len(fr), fr.ctypes.data, sys, B.shape[1], _ptr(B), _ptr(X),
                                      _stream(stream)), "gfdfd_solve_batch")
        return X

    def solve_refine_batch(self, freqs, B, sys=0, tol=1e-10, max_iter=10, out=None, stream=None):
        import torch
        X = torch.empty_like(B) if out is None else out
        fr = np.ascontiguousarray(freqs, np.int32)
        its = np.zeros(len(fr), np.int32)
        rr = np.zeros(len(fr), np.float64)
        check(lib().gfdfd_solve_refine_batch(self.h, len(fr), fr.ctypes.data, sys, B.shape[1], _ptr(B), _ptr(X), tol,
                                             max_iter, its.ctypes.data, rr.ctypes.data, _stream(stream)),
              "gfdfd_solve_refine_batch")
        return X, its.tolist(), rr.tolist()


class Workspace:
    """Factor storage reused by every call of an FWI run on one plan.

    GERMAINE refactors the impedance matrices on every objective evaluation
    (src/AC/parabolicls_AC.c:55-64 calls obj_AC per trial, src/AC/obj_AC.c:57-96
    factors each frequency) and every iteration (src/AC/fwi_FD_AC.c:229-409).
    Here the values buffer and the factor arena are allocated once, for
    nfreq frequencies in one precision; each later call assembles into the
    same buffer and refactors in place, as one hipGraph replay when graph is
    set (the capture happens on the first refactorisation).  The factors a call
    returns are overwritten by the next call that uses the workspace.
    `allocations` counts the factor arenas created (1 after any number of
    calls)."""

    def __init__(self, plan, nfreq, precision=FP64, graph=True):
        import torch
        self.plan, self.nfreq, self.precision, self.graph = plan, nfreq, precision, graph
        self.vals = torch.empty((nfreq, plan.nnz), dtype=torch.complex128, device="cuda")
        self.F = None
        self.allocations = 0

    def factors(self, physics, dh, freqs, m1, m2, **kw):
        if len(freqs) != self.nfreq:
            raise ValueError(f"workspace holds {self.nfreq} frequencies, the call has {len(freqs)}")
        self.plan.assemble(physics, dh, freqs, m1, m2, out=self.vals, **kw)
        if self.F is None:
            self.F = self.plan.factor(self.vals, self.precision)
            self.allocations += 1
        elif self.graph:
            self.F.refactor_graph(self.vals)
        else:
            self.F.refactor(self.vals)
        return self.F


def _factors(plan, ws, physics, dh, freqs, m1, m2, precision, **kw):
    """Factors of the matrices at freqs: fresh storage, or the workspace's."""
    if ws is None:
        vals = plan.assemble(physics, dh, freqs, m1, m2, **kw)
        return plan.factor(vals, precision), vals
    if ws.plan is not plan or ws.precision != precision:
        raise ValueError("workspace belongs to another plan or precision")
    return ws.factors(physics, dh, freqs, m1, m2, **kw), ws.vals


# ---- shot panels and workflows ---------------------------------------------------

def spike_panel(n, ksrc, amp, stream=None):
    import torch
    k = torch.as_tensor(np.asarray(ksrc, np.int32), device="cuda")
    B = torch.empty((len(ksrc), n), dtype=torch.complex128, device="cuda")
    check(lib().gfdfd_spike_panel(n, len(ksrc), _ptr(k), amp.real, amp.imag, _ptr(B), _stream(stream)),
          "gfdfd_spike_panel")
    return B


def sample(X, krec, stream=None):
    import torch
    nshots, n = X.shape
    k = torch.as_tensor(np.asarray(krec, np.int32), device="cuda")
    P = torch.empty((nshots, len(krec)), dtype=torch.complex64, device="cuda")
    check(lib().gfdfd_sample(n, nshots, len(krec), _ptr(k), _ptr(X), _ptr(P), _stream(stream)), "gfdfd_sample")
    return P


def adjoint_sources(n, krec, pcal, pobs, dh, stream=None):
    """calc_res_AC (MISFIT=1, STF_INV=0) and RHS_source_AC_adj for all shots:
    returns (adjoint panel [nshots, n] complex128, per-shot sum |res|^2)."""
    B, l2, _ = residual_sources(n, krec, pcal, pobs, dh, stream=stream)
    return B, l2


def residual_sources(n, krec, pcal, pobs, dh, misfit=1, stf_inv=0, rtm=False, stream=None):
    """calc_res_AC in full (src/AC/calc_res_AC.c:10-95) with RHS_source_AC_adj
    for all shots: misfit 1 (L2), 2 (logarithmic) or 3 (phase only); stf_inv=1
    estimates each shot's source-time function by Wiener deconvolution; rtm
    takes the observed data as the residual.  Returns (adjoint panel [nshots, n]
    complex128, per-shot sum |res|^2 float64, per-shot estimate complex64)."""
    import torch
    nshots, ntr = pcal.shape
    k = torch.as_tensor(np.asarray(krec, np.int32), device="cuda")
    B = torch.empty((nshots, n), dtype=torch.complex128, device="cuda")
    l2 = torch.empty(nshots, dtype=torch.float64, device="cuda")
    stf = torch.empty(nshots, dtype=torch.complex64, device="cuda")
    check(lib().gfdfd_residual_sources(n, nshots, ntr, _ptr(k), _ptr(pcal.contiguous()), _ptr(pobs.contiguous()), dh,
                                       int(misfit), int(stf_inv), int(bool(rtm)), _ptr(B), _ptr(l2), _ptr(stf),
                                       _stream(stream)), "gfdfd_residual_sources")
    return B, l2, stf


def forward_shots(plan, physics, dh, freqs, m1, m2, ksrc, krec=None, npml=0, a0_pml=90.0, s=0.0, free_surf=0,
                  precision=FP64, tol=1e-10, max_iter=10, keep_fields=False, ws=None, wavelet=None):
    """All shots x all frequencies (src/AC/forward_shot_AC.c:38-101, batched).

    Returns dict(data=[nfreq, nshots, ntr] complex64 receiver samples or None,
    fields=list of [nshots, n] complex128 panels when keep_fields, iters, relres).
    ws: a Workspace whose factor storage is reused (None: fresh storage)."""
    F, vals = _factors(plan, ws, physics, dh, freqs, m1, m2, precision, s=s, npml=npml, a0_pml=a0_pml,
                       free_surf=free_surf)
    data, fields, its, res = [], [], [], []
    for f, fr in enumerate(freqs):
        B = spike_panel(plan.n, ksrc, source_amp(physics, fr, s, dh, wavelet))
        if precision == FP64 and tol is None:
            X = F.solve(f, B)
            it, rr = 0, float("nan")
        else:
            X, it, rr = F.solve_refine(f, B, tol=tol, max_iter=max_iter)
        its.append(it)
        res.append(rr)
        if krec is not None:
            data.append(sample(X, krec))
        if keep_fields:
            fields.append(X)
    return dict(data=None if krec is None else data, fields=fields, iters=its, relres=res, factors=F, vals=vals)


def hessian_shin(plan, physics, dh, freqs, m1, m2, ksrc, npml=0, a0_pml=90.0, s=0.0, free_surf=0, precision=FP64,
                 tol=1e-10, max_iter=10, ws=None, wavelet=None):
    """Shin's pseudo-Hessian diagonal summed over frequencies and shots
    (src/AC/hessian_shin_AC.c:50-125, src/TE/hessian_shin_TE.c:53-137, batched): per frequency
    one factorisation and one panel solve for all shots, then the |J|^2 sums on the device.
    Returns AC: [n] float64; TE: (hess_sigma, hess_eps), without the reference's MAT*_NORM^2
    scaling (src/TE/hessian_shin_TE.c:147-148).  The last cell, which the reference fills from
    an unwritten entry, is 0."""
    import torch
    if physics not in (AC, TE):
        raise ValueError("SH has no Hessian path in the reference")
    n = plan.n
    F, _ = _factors(plan, ws, physics, dh, freqs, m1, m2, precision, s=s, npml=npml, a0_pml=a0_pml,
                    free_surf=free_surf)
    h1 = torch.zeros(n, dtype=torch.float64, device="cuda")
    h2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    for f, fr in enumerate(freqs):
        B = spike_panel(n, ksrc, source_amp(physics, fr, s, dh, wavelet))
        if precision == FP64 and tol is None:
            U = F.solve(f, B)
        else:
            U, _, _ = F.solve_refine(f, B, tol=tol, max_iter=max_iter)
        if physics == AC:
            o2 = omega2_grad_ac(fr, s)
            check(lib().gfdfd_hess_shin_ac(n, len(ksrc), o2.ctypes.data, _ptr(U), _ptr(m1.contiguous()),
                                           _ptr(m2.contiguous()), _ptr(h1), _stream(None)), "gfdfd_hess_shin_ac")
        else:
            o2, om = omega(TE, fr, s)
            check(lib().gfdfd_hess_shin_te(n, len(ksrc), om.ctypes.data, o2.ctypes.data, _ptr(U), _ptr(h1),
                                           _ptr(h2), _stream(None)), "gfdfd_hess_shin_te")
    return h1 if physics == AC else (h1, h2)


def hessian_gn(plan, physics, dh, freqs, m1, m2, ksrc, krec, npml=0, a0_pml=90.0, s=0.0, free_surf=0,
               precision=FP64, tol=1e-10, max_iter=10, ws=None, wavelet=None):
    """Gauss-Newton Hessian diagonal summed over frequencies, shots and receivers
    (src/AC/hessian_AC.c:50-160, src/TE/hessian_TE.c:53-177, batched): per frequency one
    factorisation, one panel solve for the shots and one for the receivers' Green's functions (a
    1/DH^2 spike at each receiver, src/AC/RHS_source_AC_hess.c:20-33,
    src/TE/RHS_source_TE_hess.c:21-32) -- shared by all shots, where the reference solves them
    again per shot -- then the pair sums on the device.  Returns AC: [n] float64; TE:
    (hess_sigma, hess_eps) without the MAT*_NORM^2 scaling.  The last cell is 0."""
    import torch
    if physics not in (AC, TE):
        raise ValueError("SH has no Hessian path in the reference")
    n = plan.n
    F, _ = _factors(plan, ws, physics, dh, freqs, m1, m2, precision, s=s, npml=npml, a0_pml=a0_pml,
                    free_surf=free_surf)
    hess = torch.zeros(n, dtype=torch.float64, device="cuda")
    h2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    for f, fr in enumerate(freqs):
        B = spike_panel(n, ksrc, source_amp(physics, fr, s, dh, wavelet))
        BG = spike_panel(n, krec, complex(np.float32(1.0) / np.float32(np.float32(dh) * np.float32(dh))))
        if precision == FP64 and tol is None:
            U, GR = F.solve(f, B), F.solve(f, BG)
        else:
            U, _, _ = F.solve_refine(f, B, tol=tol, max_iter=max_iter)
            GR, _, _ = F.solve_refine(f, BG, tol=tol, max_iter=max_iter)
        if physics == AC:
            o2 = omega2_grad_ac(fr, s)
            check(lib().gfdfd_hess_gn_ac(n, len(ksrc), len(krec), o2.ctypes.data, _ptr(U), _ptr(GR),
                                         _ptr(m1.contiguous()), _ptr(m2.contiguous()), _ptr(hess), _stream(None)),
                  "gfdfd_hess_gn_ac")
        else:
            o2, om = omega(TE, fr, s)
            check(lib().gfdfd_hess_gn_te(n, len(ksrc), len(krec), om.ctypes.data, o2.ctypes.data, _ptr(U),
                                         _ptr(GR), _ptr(hess), _ptr(h2), _stream(None)), "gfdfd_hess_gn_te")
    return hess if physics == AC else (hess, h2)


def apply_hess(grad, hess, eps_hess, stream=None):
    """Diagonal-Hessian preconditioning of a gradient in place (src/AC/apply_hess_AC.c:20-37):
    grad *= 1 / (hess + eps_hess * max|hess|); both float64 cuda tensors of n cells."""
    check(lib().gfdfd_apply_hess(grad.numel(), float(eps_hess), _ptr(hess.contiguous()), _ptr(grad), _stream(stream)),
          "gfdfd_apply_hess")
    return grad


def model_update(alpha, direction, m_old, lo, hi, out=None, stream=None):
    """Trial model of a line-search step (src/calc_mat_change_wolfe.c:27-46):
    m_old + alpha * direction in float, where a value below lo (MAT1_LOW), above
    hi (MAT1_UP) or below zero keeps m_old.  m_old float32 cuda [ny, nx];
    direction float64 cuda of as many cells.  Returns out (a new tensor if None)."""
    import torch
    m_old = m_old.contiguous()
    if direction.numel() != m_old.numel():
        raise ValueError("model_update: direction and model differ in size")
    out = torch.empty_like(m_old) if out is None else out
    check(lib().gfdfd_model_update(m_old.numel(), float(alpha), _ptr(direction.contiguous()), _ptr(m_old),
                                   float(lo), float(hi), _ptr(out), _stream(stream)), "gfdfd_model_update")
    return out


def model_update_norm(alpha, direction, m_old, norm, lo, hi, out=None, stream=None):
    """The TE line search's trial model (src/TE/parabolicls_TE.c:30-31,64-69):
    the normalised material (1/norm) * m_old stepped by alpha * direction,
    clamped to [lo/norm, hi/norm] and scaled back by norm, in float."""
    import torch
    m_old = m_old.contiguous()
    if direction.numel() != m_old.numel():
        raise ValueError("model_update_norm: direction and model differ in size")
    out = torch.empty_like(m_old) if out is None else out
    check(lib().gfdfd_model_update_norm(m_old.numel(), float(alpha), _ptr(direction.contiguous()), _ptr(m_old),
                                        float(norm), float(lo), float(hi), _ptr(out), _stream(stream)),
          "gfdfd_model_update_norm")
    return out


def objective(plan, physics, dh, freqs, m1, m2, ksrc, krec, pobs, npml=0, a0_pml=90.0, s=0.0, free_surf=0,
              precision=FP64, tol=1e-10, max_iter=10, misfit=1, stf_inv=0, ws=None, wavelet=None):
    """Misfit only (obj_AC, src/AC/obj_AC.c:46-135): forward panel per frequency,
    receiver samples, residual norms (calc_res_AC with misfit / stf_inv);
    0.5 * sum |res|^2 over frequencies and shots.  The plan (ordering, symbolic
    tree) is reused, and with a Workspace the factor storage too, as every
    line-search trial of an FWI iteration does."""
    n = plan.n
    F, _ = _factors(plan, ws, physics, dh, freqs, m1, m2, precision, s=s, npml=npml, a0_pml=a0_pml,
                    free_surf=free_surf)
    l2 = 0.0
    for f, fr in enumerate(freqs):
        B = spike_panel(n, ksrc, source_amp(physics, fr, s, dh, wavelet))
        if precision == FP64 and tol is None:
            U = F.solve(f, B)
        else:
            U, _, _ = F.solve_refine(f, B, tol=tol, max_iter=max_iter)
        _, l2s, _ = residual_sources(n, krec, sample(U, krec), pobs[f], dh, misfit=misfit, stf_inv=stf_inv)
        l2 += float(l2s.sum())
    return 0.5 * l2


def initial_step(direction, m, eps_scale):
    """First trial step of the parabolic line search (src/AC/parabolicls_AC.c:35-37):
    EPS_SCALE * max(vp) / max(Hgrad), signed maxima as maximum_m takes them."""
    return np.float32(eps_scale) * np.float32(float(m.max())) / np.float32(float(direction.max()))


def opt_step(L2t, epst):
    """Step length at the vertex of the parabola through (epst[i], L2t[i]), i = 0..2
    (calc_opt_step, src/calc_opt_step.c:56-102), clamped as the reference clamps it:
    a concave fit with L2t[1] > L2t[2] takes epst[2]; a vertex beyond epst[2] takes
    epst[2]; a negative one takes epst[0]."""
    e = np.asarray(epst, np.float64)
    A = np.stack([np.ones(3), e, e * e], axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.linalg.solve(A, np.asarray(L2t, np.float64))
        opteps = np.float64(-x[1]) / np.float64(2.0 * x[2])
    if 2.0 * x[2] < 0.0 and L2t[1] > L2t[2]:
        opteps = e[2]
    if opteps > e[2]:
        opteps = e[2]
    if opteps < 0.0:
        opteps = e[0]
    return float(opteps)


def parabolic_line_search(obj, L2, alpha, scalefac=2.0, stepmax=6):
    """Parabolic line search (parabolicls_AC, src/AC/parabolicls_AC.c:28-153).

    obj(alpha) returns the misfit of the model stepped by alpha; L2 is the misfit
    at alpha = 0.  The search halves the step (by scalefac) until the misfit
    falls, then widens it until the misfit rises again, and fits a parabola
    through the three points.  Returns (step, trials), step 0.0 when no step
    within stepmax reductions lowers the misfit; trials lists (alpha, misfit)."""
    f32 = np.float32
    alpha, scalefac = f32(alpha), f32(scalefac)
    L2t = [None, L2, 0.0, 0.0]          # 1-based, as the reference indexes them
    epst = [None, 0.0, 0.0, 0.0]        # epst[1] stays 0: the step of L2t[1]
    step1 = step2 = countstep = 0
    itests = iteste = 2
    trials = []
    while step2 != 1 or step1 != 1:
        for itest in range(itests, iteste + 1):
            L2t[itest] = obj(float(alpha))
            trials.append((float(alpha), L2t[itest]))
        if step1 == 0 and L2t[1] <= L2t[2]:
            alpha = alpha / scalefac
            countstep += 1
        if step1 == 1 and L2t[2] < L2t[3]:
            epst[3] = float(alpha)
            step2 = 1
        if step1 == 1 and L2t[2] >= L2t[3]:
            epst[3] = float(alpha)
            alpha = alpha + alpha / scalefac
            countstep += 1
        if step1 == 0 and L2t[1] > L2t[2]:
            epst[2] = float(alpha)
            step1 = 1
            itests = iteste = 3
            countstep = 0
            alpha = alpha + alpha / scalefac
        if step1 == 0 and countstep > stepmax:
            return 0.0, trials
        if step1 == 1 and countstep > stepmax:
            step2 = 1
    return opt_step(L2t[1:], epst[1:]), trials


def line_search(plan, dh, freqs, vp, rho, direction, ksrc, krec, pobs, L2, eps_scale, lo, hi, scalefac=2.0,
                stepmax=6, ws=None, alpha0=None, **kw):
    """One FWI iteration's step search for the acoustic model (src/AC/fwi_FD_AC.c
    calling parabolicls_AC): each trial steps vp along direction on the device
    (model_update into one reused buffer) and evaluates objective() on the same
    plan and factor storage (ws, or one workspace for the whole search).
    The first trial step is EPS_SCALE * max(vp) / max(direction) with signed
    maxima (src/AC/parabolicls_AC.c:35-37), or alpha0 when given (the signed
    rule degenerates for a direction whose entries are almost all negative).
    Returns (step, trials); the caller applies model_update(step, ...)."""
    if ws is None:
        ws = Workspace(plan, len(freqs), kw.get("precision", FP64))
    trial = model_update(0.0, direction, vp, lo, hi)
    def obj(alpha):
        model_update(alpha, direction, vp, lo, hi, out=trial)
        return objective(plan, AC, dh, freqs, trial, rho, ksrc, krec, pobs, ws=ws, **kw)
    a0 = initial_step(direction, vp, eps_scale) if alpha0 is None else alpha0
    return parabolic_line_search(obj, L2, a0, scalefac, stepmax)


class LBFGS:
    """Quasi-Newton search direction of the FWI loop (GRAD_METHOD=2): the two-loop
    recursion of src/LBFGS.c:36-160 over the last `memory` (s, y) pairs, s the
    model change and y the gradient change of consecutive iterations, newest
    to oldest then back, the initial inverse Hessian gamma = s.y / y.y of the
    newest pair; the first call returns -grad (:162-170).  The memory drops its
    oldest pair once full (rot_LBFGS_vec, src/AC/fwi_FD_AC.c:231-237), and a
    direction that is not a descent direction clears it (check_descent,
    src/check_descent.c:19-40).  O(n) vector work per FWI iteration, run on
    the host in float64 (the reference runs it on MPI rank 0 in float)."""

    def __init__(self, memory=5):
        self.memory = memory
        self.pairs = []          # (s, y), oldest first
        self.prev = None

    def direction(self, grad, model):
        import torch
        g = grad.detach().double().cpu().numpy().ravel()
        x = model.detach().double().cpu().numpy().ravel()
        if self.prev is not None:
            self.pairs.append((x - self.prev[1], g - self.prev[0]))
            if len(self.pairs) > self.memory:
                self.pairs.pop(0)
        self.prev = (g.copy(), x.copy())
        q = -g
        if self.pairs:
            rho = [1.0 / float(y @ s_) if abs(float(y @ s_)) > 0.0 else 0.0 for s_, y in self.pairs]
            alpha = [0.0] * len(self.pairs)
            for k in range(len(self.pairs) - 1, -1, -1):
                s_, y = self.pairs[k]
                alpha[k] = rho[k] * float(s_ @ q)
                q = q - alpha[k] * y
            s_n, y_n = self.pairs[-1]
            r = (float(y_n @ s_n) / float(y_n @ y_n)) * q
            for k, (s_, y) in enumerate(self.pairs):
                r = r + s_ * (alpha[k] - rho[k] * float(y @ r))
            q = r
        if -float(q @ g) / float(g @ g) < 0.0:
            self.pairs.clear()
        return torch.as_tensor(q.reshape(grad.shape), device=grad.device)


def tikhonov_te(sigma, eps, dh, beta, norm, lam, grad_sigma=None, grad_eps=None, stream=None):
    """Tikhonov regularisation of the TE materials (Tikhonov_cost_TE,
    src/TE/Tikhonov_cost_TE.c:21-45, and Tikhonov_grad_TE, src/TE/Tikhonov_grad_TE.c:20-37):
    beta, norm, lam are (sigma, epsilon) pairs of BETA_MAT*, MAT*_NORM, LAMBDA_*.
    Adds LAMBDA * BETA * Laplacian to the gradients that are given and returns
    the two costs (tikh_sigma, tikh_epsilon); the misfit the reference compares
    is L2 + lam[0] * tikh_sigma + lam[1] * tikh_epsilon when a lam is positive."""
    import torch
    ny, nx = sigma.shape
    cost = torch.empty(2, dtype=torch.float64, device="cuda")
    gp = lambda g: None if g is None else _ptr(g)
    check(lib().gfdfd_tikhonov_te(nx, ny, float(dh), float(beta[0]), float(beta[1]), float(norm[0]), float(norm[1]),
                                  float(lam[0]), float(lam[1]), _ptr(sigma.contiguous()), _ptr(eps.contiguous()),
                                  gp(grad_sigma), gp(grad_eps), _ptr(cost), _stream(stream)), "gfdfd_tikhonov_te")
    c = cost.cpu().numpy()
    return float(c[0]), float(c[1])


def line_search_te(plan, dh, freqs, eps, sigma, dir_eps, dir_sigma, ksrc, krec, pobs, L2, eps_scale, norm,
                   low, up, scalefac=2.0, stepmax=6, tikhonov=None, ws=None, alpha0=None, **kw):
    """One FWI iteration's step search for the TE model (parabolicls_TE,
    src/TE/parabolicls_TE.c:28-160): both materials step by the same alpha on
    their normalised scales (norm, low, up: (MAT1 for sigma, MAT2 for epsilon)
    pairs; dir_* are the search directions of the normalised materials), the
    first trial step is EPS_SCALE * max(sigma/MAT1_NORM) / max(dir_sigma), and
    each trial's misfit comes from objective() on the same plan, plus the
    Tikhonov term (Tikhonov_cost_TE) when tikhonov = (beta, lam) pairs are
    given and a lam is positive.  L2 is the start model's misfit with that
    term.  alpha0, when given, replaces the signed-maximum first step.
    Returns (step, trials)."""
    f32 = np.float32
    if not (norm[0] > 0 and norm[1] > 0):
        raise ValueError("line_search_te: the normalisations must be positive")
    if ws is None:
        ws = Workspace(plan, len(freqs), kw.get("precision", FP64))
    ts = model_update_norm(0.0, dir_sigma, sigma, norm[0], low[0], up[0])
    te = model_update_norm(0.0, dir_eps, eps, norm[1], low[1], up[1])
    def obj(alpha):
        model_update_norm(alpha, dir_sigma, sigma, norm[0], low[0], up[0], out=ts)
        model_update_norm(alpha, dir_eps, eps, norm[1], low[1], up[1], out=te)
        l2 = objective(plan, TE, dh, freqs, te, ts, ksrc, krec, pobs, ws=ws, **kw)
        if tikhonov is not None and (tikhonov[1][0] > 0 or tikhonov[1][1] > 0):
            cs, ce = tikhonov_te(ts, te, dh, tikhonov[0], norm, tikhonov[1])
            l2 += tikhonov[1][0] * cs + tikhonov[1][1] * ce
        return l2
    # max(a * sigma) = a * max(sigma) for a > 0: float multiplication is monotone
    if alpha0 is None:
        alpha0 = f32(eps_scale) * (f32(1.0 / norm[0]) * f32(float(sigma.max()))) / f32(float(dir_sigma.max()))
    return parabolic_line_search(obj, L2, alpha0, scalefac, stepmax)


def precond(grad, nx, ny, dh, vref, fc_low, filt_size_x, filt_size_y, spatfilter=1, nzero_rows=0, stream=None):
    """precond (src/precond.c:17-26) without the taper file: Gaussian smoothing
    over filt_size_x / filt_size_y wavelengths vref / fc_low (gauss_filt), then
    the first nzero_rows rows zeroed (taper_grad_hor with GRADT2 + FSSHIFT).
    grad: float64 cuda of nx * ny cells; returns a new tensor."""
    import torch
    grad = grad.contiguous()
    if grad.numel() != nx * ny:
        raise ValueError("precond: field size differs from nx * ny")
    out = torch.empty_like(grad)
    check(lib().gfdfd_precond(nx, ny, float(dh), float(vref), float(fc_low), float(filt_size_x), float(filt_size_y),
                              int(spatfilter), int(nzero_rows), _ptr(grad), _ptr(out), _stream(stream)),
          "gfdfd_precond")
    return out


def taper_file(grad, taper, npml, free_surf, stream=None):
    """taper_grad_file (src/taper_grad_file.c:22-66): grad (float64 cuda of the
    grid with PML) times taper (float32 cuda [ny0, nx0], taper.bin as
    formats.read_model_file reads it) on the model part, PML refilled by the
    reference's edge copies.  Returns a new tensor."""
    import torch
    ny0, nx0 = taper.shape
    fs = 0 if free_surf else npml
    if grad.numel() != (nx0 + 2 * npml) * (ny0 + npml + fs):
        raise ValueError("taper_file: field size differs from the grid with PML")
    grad = grad.contiguous()
    out = torch.empty_like(grad)
    check(lib().gfdfd_taper_file(nx0, ny0, npml, int(free_surf), _ptr(taper.contiguous()), _ptr(grad), _ptr(out),
                                 _stream(stream)), "gfdfd_taper_file")
    return out


def grad_frame(grad, nx0, ny0, npml, free_surf, stream=None):
    """cp_grad_frame (src/cp_grad_frame.c:102-208): damp the gradient (float64
    cuda, grid with PML) in the PML with the reference's taper, in place."""
    fs = 0 if free_surf else npml
    if grad.numel() != (nx0 + 2 * npml) * (ny0 + npml + fs) or not grad.is_contiguous():
        raise ValueError("grad_frame: the gradient must be the contiguous grid with PML")
    check(lib().gfdfd_grad_frame(nx0, ny0, npml, int(free_surf), _ptr(grad), _stream(stream)), "gfdfd_grad_frame")
    return grad


class ShotTaper:
    """ass_grad_AC/TE's per-shot gradient taper, SWS_TAPER_CIRCULAR_PER_SHOT
    (src/AC/ass_grad_AC.c:37-43, src/TE/ass_grad_TE.c:40-49).

    mode 1 (taper_grad_shot, src/taper_grad_shot.c:34-87): every shot's gradient is
    multiplied by f0 + (1 - f0)(0.5 - 0.5 cos(pi r/SRTRADIUS))^a within SRTRADIUS of the
    shot's source and of every receiver; mode 2 (taper_grad_shot1, src/taper_grad_shot1.c):
    it is 1e-5 on the (2 FILTSIZE + 1)^2 boxes around every source and receiver.
    Distances and boxes use the reference's coordinates: cell (ix, iy) of the
    PML-extended grid at (ix DH, iy DH) against srcpos in metres and recpos in grid
    indices, as sources() and receiver() return them (par.read_sources /
    par.read_receivers), without the PML offset the solves add.
    srcpos: [nsrc, 2] metres; recpos: [ntr, 2] grid indices; shots: the rows of srcpos
    that the gradient's shot panel holds, in panel order (default: all)."""

    def __init__(self, mode, nx, dh, srcpos, recpos, srtradius=0.0, filtsize=0, shots=None):
        import torch
        if mode not in (1, 2):
            raise ValueError("SWS_TAPER_CIRCULAR_PER_SHOT must be 1 or 2")
        srcpos = np.asarray(srcpos, np.float32).reshape(-1, 2)
        recpos = np.asarray(recpos, np.int32).reshape(-1, 2)
        shots = np.arange(len(srcpos)) if shots is None else np.asarray(shots)
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
        self.mode, self.nx, self.dh = int(mode), int(nx), float(dh)
        self.srtradius, self.filtsize = float(srtradius), int(filtsize)
        self.nshots = len(shots)
        self._srcx, self._srcy = dev(srcpos[shots, 0]), dev(srcpos[shots, 1])
        # iround(srcpos / DH): the float quotient, floor(x + 0.5) in double
        q = srcpos / np.float32(dh)
        self._si = dev(np.floor(q[:, 0].astype(np.float64) + 0.5).astype(np.int32))
        self._sj = dev(np.floor(q[:, 1].astype(np.float64) + 0.5).astype(np.int32))
        self._ri, self._rj = dev(recpos[:, 0]), dev(recpos[:, 1])

    def device_struct(self, nshots):
        if nshots != self.nshots:
            raise ValueError(f"the taper holds {self.nshots} shots, the panel {nshots}")
        return ShotTaperC(self.mode, self.nx, self.dh, self.srtradius, self.filtsize, self._srcx.data_ptr(),
                          self._srcy.data_ptr(), len(self._si), self._si.data_ptr(), self._sj.data_ptr(),
                          len(self._ri), self._ri.data_ptr(), self._rj.data_ptr())


def grad_obj(plan, physics, dh, freqs, m1, m2, ksrc, krec, pobs, npml=0, a0_pml=90.0, s=0.0, free_surf=0,
             precision=FP64, tol=1e-10, max_iter=10, rtm=False, misfit=1, stf_inv=0, ws=None, wavelet=None, taper=None):
    """Misfit and gradient summed over frequencies and shots (src/AC/grad_obj_AC.c:50-168,
    src/TE/grad_obj_TE.c:53-172).  pobs: list per frequency of complex64 cuda [nshots, ntr].
    Returns (L2 = 0.5 * sum |res|^2, grad (AC: [n] float64; TE: (grad_sigma, grad_eps))).

    misfit (MISFIT 1/2/3) and stf_inv (STF_INV) select calc_res_AC's residual
    and the Wiener source-time-function estimate, whose phase weights each
    shot's gradient (src/AC/ass_grad_AC.c:30-31, src/TE/ass_grad_TE.c:31-35).
    rtm=True is the reverse-time-migration pass (RTM_FD_AC, src/AC/RTM_FD_AC.c:190):
    the residual is the observed data itself (src/AC/calc_res_AC.c:77), with the
    same mute, so the returned gradient is the image before precond().
    ws: a Workspace whose factor storage is reused (None: fresh storage).
    wavelet: (samples, ndt) of an SU source wavelet (READ_WAVELET = 1), see source_amp.
    taper: ass_grad's per-shot taper (SWS_TAPER_CIRCULAR_PER_SHOT), a ShotTaper."""
    import torch
    n = plan.n
    F, _ = _factors(plan, ws, physics, dh, freqs, m1, m2, precision, s=s, npml=npml, a0_pml=a0_pml,
                    free_surf=free_surf)
    tp = None if taper is None else taper.device_struct(len(ksrc))
    g1 = torch.zeros(n, dtype=torch.float64, device="cuda")
    g2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    l2 = 0.0
    adj_sys = 0 if physics == AC else 2      # AC solves with M (grad_obj_AC.c:147), TE with M^T (grad_obj_TE.c:149)
    k = torch.as_tensor(np.asarray(krec, np.int32), device="cuda")
    for f, fr in enumerate(freqs):
        B = spike_panel(n, ksrc, source_amp(physics, fr, s, dh, wavelet))
        if precision == FP64 and tol is None:
            U = F.solve(f, B)
        else:
            U, _, _ = F.solve_refine(f, B, tol=tol, max_iter=max_iter)
        pcal = sample(U, krec)
        BA, l2s, stf = residual_sources(n, krec, pcal, pobs[f], dh, misfit=misfit, stf_inv=stf_inv, rtm=rtm)
        l2 += float(l2s.sum())
        if precision == FP64 and tol is None:
            LAM = F.solve(f, BA, sys=adj_sys)
        else:
            LAM, _, _ = F.solve_refine(f, BA, sys=adj_sys, tol=tol, max_iter=max_iter)
        nshots = len(ksrc)
        tpp = None if tp is None else C.byref(tp)
        if physics == AC:
            o2 = omega2_grad_ac(fr, s)
            check(lib().gfdfd_grad_ac_taper(n, nshots, o2.ctypes.data, _ptr(U), _ptr(LAM), _ptr(m1.contiguous()),
                                            _ptr(m2.contiguous()), _ptr(stf), tpp, _ptr(g1), _stream(None)),
                  "gfdfd_grad_ac_taper")
        elif physics == TE:
            o2, om = omega(TE, fr, s)
            check(lib().gfdfd_grad_te_taper(n, nshots, om.ctypes.data, o2.ctypes.data, _ptr(U), _ptr(LAM), _ptr(stf),
                                            tpp, _ptr(g1), _ptr(g2), _stream(None)), "gfdfd_grad_te_taper")
        else:
            raise ValueError("SH has no gradient path in the reference (src/germaine.c:112-120)")
    grad = g1 if physics == AC else (g1, g2)
    return 0.5 * l2, grad
