"""GPU parity of the FWI gradient path against the oracle.

grad_obj on the device (batched assembly and factorisation, forward solves,
receiver sampling, residual with the reference's signed mute, adjoint
sources conj(res)/DH^2, adjoint solves with M for AC and M^T for TE, and the
gradient kernels) against the oracle's per-shot restatement of
src/AC/grad_obj_AC.c:50-168 and src/TE/grad_obj_TE.c:53-172 with FP64 band-LU
solves.  The reference accumulates the gradient and the misfit in float,
the device in double, so the bar is a relative error of 1e-5 on the gradient
(FP32 rounding of the summands) and on the misfit -- in FP64 and with FP32
factors refined in FP64 to a relative residual of 1e-10.
"""
import sys

import numpy as np
import pytest

import germaine_amd as G
from germaine_amd import fdfd
from oracle import oracle as O
from tests._grids import smooth_model

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NX, NY, NPML = 52, 38, 8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def models(phys):
    j, i = np.mgrid[0:NY, 0:NX]
    bump = np.exp(-(((i - 26) / 6.0) ** 2 + ((j - 22) / 5.0) ** 2))
    if phys == G.AC:
        m1, m2 = smooth_model(NX, NY, seed=13)
        true1 = (m1 + 180.0 * bump).astype(np.float32)
        return dict(dh=10.0, freqs=[6.0, 11.5], m1=m1, m2=m2, t1=true1, t2=m2, adj=0)
    eps0 = 8.854187817e-12
    rng = np.random.default_rng(3)
    m1 = ((5.0 + rng.random((NY, NX))) * eps0).astype(np.float32)
    m2 = (0.002 + 0.002 * rng.random((NY, NX))).astype(np.float32)
    true1 = (m1 + 2.0 * eps0 * bump).astype(np.float32)
    return dict(dh=0.04, freqs=[1.2e8, 2.6e8], m1=m1, m2=m2, t1=true1, t2=m2, adj=2)


def oracle_csc(phys, dh, f, m1, m2):
    px, py = O.pml_profiles(NX, NY, NPML, dh, 90.0, f, 0.0, 0)
    return O.csc_of_M(NX, NY, *O.assemble(phys, NX, NY, dh, f, 0.0, m1, m2, px, py))


def oracle_fields(phys, dh, f, m1, m2, srcs, rhs=None, sysno=0):
    csc = oracle_csc(phys, dh, f, m1, m2)
    n = NX * NY
    if rhs is None:
        rhs = np.zeros((n, len(srcs)), complex)
        for s, k in enumerate(srcs):
            rhs[k, s] = O.source_amp(phys, f, 0.0, dh)
    X, info = O.band_solve(*csc, rhs, sysno, O.grid_transpose_perm(NX, NY))
    assert info == 0
    return csc, X


@pytest.mark.parametrize("rtm", [False, True])
@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
@pytest.mark.parametrize("phys", [G.AC, G.TE])
def test_gradient_matches_oracle(phys, precision, rtm):
    """rtm: the migration pass (src/AC/RTM_FD_AC.c:190), whose residual is the
    observed data (src/AC/calc_res_AC.c:77): the oracle's residual with zero
    predicted data."""
    say("gradient", phys, precision, "rtm" if rtm else "fwi")
    m = models(phys)
    dh, freqs, n = m["dh"], m["freqs"], NX * NY
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    # observed data from the true model, with two muted samples per frequency:
    # one zero, one with both parts negative (the signed mute, src/AC/calc_res_AC.c:82-85)
    pobs = []
    for f in freqs:
        _, Ut = oracle_fields(phys, dh, f, m["t1"], m["t2"], srcs)
        p = Ut[recs, :].T.astype(np.complex64)
        p[0, 3] = 0
        p[2, 7] = -abs(p[2, 7].real) - 1e-3j * abs(p[2, 7])
        pobs.append(p)

    pl = fdfd.Plan(NX, NY)
    tol = None if precision == G.FP64 else 1e-10
    l2_dev, grad = fdfd.grad_obj(pl, phys, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, recs,
                                 [dev(p) for p in pobs], npml=NPML, precision=precision, tol=tol, rtm=rtm)
    torch.cuda.synchronize()

    g1 = np.zeros(n, np.float32)
    g2 = np.zeros(n, np.float32)
    l2 = 0.0
    for fi, f in enumerate(freqs):
        csc, U = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs)
        rhs = np.zeros((n, len(srcs)), complex)
        for s in range(len(srcs)):
            pcal = np.zeros(len(recs), np.complex64) if rtm else U[recs, s].astype(np.complex64)
            res, l2s = O.residual(pcal, pobs[fi][s])
            assert res[3] == 0 or s != 0
            l2 += l2s
            rhs[:, s] = O.adjoint_rhs(n, recs, res, dh)
        _, LAM = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs, rhs, m["adj"])
        for s in range(len(srcs)):
            # vec2mat's one-cell shift reads x[n], which no solve writes: the
            # device leaves that cell unchanged, the oracle adds zero there
            if phys == G.AC:
                O.grad_ac_accumulate(NX, NY, f, 0.0, U[:, s], 0j, LAM[:, s], 0j, m["m1"], m["m2"], g1)
            else:
                O.grad_te_accumulate(NX, NY, f, 0.0, U[:, s], 0j, LAM[:, s], 0j, g1, g2)
    l2 *= 0.5

    def relerr(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    assert abs(l2_dev - l2) <= 1e-5 * abs(l2), (l2_dev, l2)
    if rtm:
        # the misfit of the migration pass is the data's energy, muted samples included
        assert abs(l2 - 0.5 * sum(float(np.sum(np.abs(p.astype(np.complex128)) ** 2)) for p in pobs)) <= 1e-5 * l2
    else:
        # the line search's misfit-only pass (obj_AC) on the same plan
        l2_obj = fdfd.objective(pl, phys, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, recs,
                                [dev(p) for p in pobs], npml=NPML, precision=precision, tol=tol)
        assert abs(l2_obj - l2) <= 1e-5 * abs(l2), (l2_obj, l2)
    if phys == G.AC:
        e = relerr(grad.cpu().numpy(), g1.astype(np.float64))
        say("  misfit", l2_dev, "gradient rel err", e)
        assert e <= 1e-5
    else:
        es = relerr(grad[0].cpu().numpy(), g1.astype(np.float64))
        ee = relerr(grad[1].cpu().numpy(), g2.astype(np.float64))
        say("  misfit", l2_dev, "gradient rel err sigma", es, "eps", ee)
        assert es <= 1e-5 and ee <= 1e-5


@pytest.mark.parametrize("misfit,stf_inv,rtm", [(1, 0, 0), (2, 0, 0), (3, 0, 0), (1, 1, 0), (2, 1, 0), (1, 0, 1)])
def test_residual_sources_match_oracle(misfit, stf_inv, rtm):
    """calc_res_AC's branches (src/AC/calc_res_AC.c:21-88) with RHS_source_AC_adj on
    the device against the oracle, per shot: the Wiener estimate in the
    reference's summation order and complex float rounding (bit-exact), the
    residuals (misfit 1 and rtm bit-exact; misfits 2 and 3 within four float
    spacings of the largest operand, |log| or pi: clogf/cargf are rounded once
    from double here, by glibc's own methods in the oracle), the adjoint panel
    and the misfit (double sums here, float in the reference)."""
    say("residual sources", misfit, stf_inv, rtm)
    rng = np.random.default_rng(100 + 10 * misfit + stf_inv)
    n, nshots, ntr = 500, 7, 60
    krec = rng.choice(n, ntr, replace=False).astype(np.int32)
    krec[5] = krec[9]                                  # a duplicate receiver: the last one wins
    pcal = ((rng.standard_normal((nshots, ntr)) + 1j * rng.standard_normal((nshots, ntr))) * 1e-3).astype(np.complex64)
    w = 0.7 * np.exp(0.9j)
    pobs = (w * pcal + 1e-4 * (rng.standard_normal((nshots, ntr)) + 1j * rng.standard_normal((nshots, ntr))))
    pobs = pobs.astype(np.complex64)
    pobs[1, 4] = 1e-30           # muted (both parts below 1e-20), with a finite logarithm
    dh = 12.5
    B, l2, stf = fdfd.residual_sources(n, krec, dev(pcal), dev(pobs), dh, misfit=misfit, stf_inv=stf_inv, rtm=rtm)
    B, l2, stf = B.cpu().numpy(), l2.cpu().numpy(), stf.cpu().numpy()
    for s_ in range(nshots):
        res, l2s, st = O.residual_ex(pcal[s_], pobs[s_], misfit=misfit, stf_inv=stf_inv, rtm=rtm)
        assert stf[s_] == np.complex64(st), (s_, stf[s_], st)
        want = O.adjoint_rhs(n, krec, res, dh)
        if misfit in (2, 3) and not rtm:
            # a few float spacings of the largest operand (|log| or pi) of the unmuted receivers
            live = ~((pobs[s_].real < 1e-20) & (pobs[s_].imag < 1e-20))
            big = max(np.pi, float(np.max(np.abs(np.log(np.abs(pobs[s_][live]))))),
                      float(np.max(np.abs(np.log(np.abs(pcal[s_]))))))
            tol = 4 * float(np.spacing(np.float32(big))) / (dh * dh)
            assert np.max(np.abs(B[s_] - want)) <= tol, (s_, np.max(np.abs(B[s_] - want)), tol)
        else:
            assert np.array_equal(B[s_], want), s_
        assert abs(l2[s_] - l2s) <= 1e-5 * abs(l2s)


@pytest.mark.parametrize("misfit,stf_inv", [(2, 0), (3, 0), (1, 1)])
@pytest.mark.parametrize("phys", [G.AC, G.TE])
def test_gradient_misfits_and_stf_match_oracle(phys, misfit, stf_inv):
    """grad_obj with the logarithmic and phase-only misfits and with the Wiener
    source-time-function estimate (src/AC/calc_res_AC.c:21-74), whose phase
    weights each shot's gradient (src/AC/ass_grad_AC.c:30-31,
    src/TE/ass_grad_TE.c:31-35), against the oracle's per-shot restatement
    with band-LU solves: relative 1e-5 as the L2 gradient test.  The observed
    data carry a complex source factor so the estimate is not trivial."""
    say("gradient misfit", phys, misfit, "stf", stf_inv)
    m = models(phys)
    dh, freqs, n = m["dh"], m["freqs"], NX * NY
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    pobs = []
    for f in freqs:
        _, Ut = oracle_fields(phys, dh, f, m["t1"], m["t2"], srcs)
        pobs.append((0.6 * np.exp(-0.5j) * Ut[recs, :].T).astype(np.complex64))
    pl = fdfd.Plan(NX, NY)
    l2_dev, grad = fdfd.grad_obj(pl, phys, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, recs,
                                 [dev(p) for p in pobs], npml=NPML, precision=G.FP64, tol=None, misfit=misfit,
                                 stf_inv=stf_inv)
    g1 = np.zeros(n, np.float32)
    g2 = np.zeros(n, np.float32)
    l2 = 0.0
    for fi, f in enumerate(freqs):
        csc, U = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs)
        rhs = np.zeros((n, len(srcs)), complex)
        stfs = []
        for s_ in range(len(srcs)):
            res, l2s, st = O.residual_ex(U[recs, s_].astype(np.complex64), pobs[fi][s_], misfit=misfit,
                                         stf_inv=stf_inv)
            stfs.append(st)
            l2 += l2s
            rhs[:, s_] = O.adjoint_rhs(n, recs, res, dh)
        _, LAM = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs, rhs, m["adj"])
        for s_ in range(len(srcs)):
            if phys == G.AC:
                O.grad_ac_accumulate(NX, NY, f, 0.0, U[:, s_], 0j, LAM[:, s_], 0j, m["m1"], m["m2"], g1, stf=stfs[s_])
            else:
                O.grad_te_accumulate(NX, NY, f, 0.0, U[:, s_], 0j, LAM[:, s_], 0j, g1, g2, stf=stfs[s_])
        if stf_inv:
            assert all(abs(st - 1) > 0.1 for st in stfs)
    l2 *= 0.5

    def relerr(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    assert abs(l2_dev - l2) <= 1e-5 * abs(l2), (l2_dev, l2)
    if phys == G.AC:
        e = relerr(grad.cpu().numpy(), g1.astype(np.float64))
        say("  misfit", l2_dev, "gradient rel err", e)
        assert e <= 1e-5
    else:
        es = relerr(grad[0].cpu().numpy(), g1.astype(np.float64))
        ee = relerr(grad[1].cpu().numpy(), g2.astype(np.float64))
        say("  misfit", l2_dev, "gradient rel err sigma", es, "eps", ee)
        assert es <= 1e-5 and ee <= 1e-5


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
@pytest.mark.parametrize("phys", [G.AC, G.TE])
def test_shin_hessian_matches_oracle(phys, precision):
    """hessian_shin (src/AC/hessian_shin_AC.c:50-125, src/TE/hessian_shin_TE.c:53-137)
    on the device against the oracle's per-shot float restatement with FP64
    band-LU fields: relative 1e-5 (the reference accumulates in float, the
    device in double).  TE at one frequency: there |Omega^2 u|^2 exceeds the
    float range on almost every cell (the reference's sum is inf there; the
    device's double sum is not), so the epsilon Hessian is compared on the
    cells where the float restatement is finite and checked everywhere as
    |Omega|^2 times the sigma Hessian, which the oracle pins."""
    say("shin hessian", phys, precision)
    m = models(phys)
    dh, freqs, n = m["dh"], m["freqs"], NX * NY
    if phys == G.TE:
        freqs = freqs[:1]
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    pl = fdfd.Plan(NX, NY)
    out = fdfd.hessian_shin(pl, phys, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, npml=NPML, precision=precision,
                            tol=None if precision == G.FP64 else 1e-10)
    hs = [out] if phys == G.AC else list(out)
    refs = [np.zeros(n, np.float32) for _ in hs]
    for f in freqs:
        _, U = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs)
        for s in range(len(srcs)):
            if phys == G.AC:
                O.hess_shin_ac_accumulate(NX, NY, f, 0.0, U[:, s], 0j, m["m1"], m["m2"], refs[0])
            else:
                O.hess_shin_te_accumulate(NX, NY, f, 0.0, U[:, s], 0j, refs[0], refs[1])
    for i, (h, ref) in enumerate(zip(hs, refs)):
        h = h.cpu().numpy()
        fin = np.isfinite(ref[:-1])
        assert fin.all() if i == 0 else fin.any()
        err = float(np.max(np.abs(h[:-1][fin] - ref[:-1][fin])) / np.max(np.abs(ref[:-1][fin])))
        say("  rel err", err, "on", int(fin.sum()), "cells")
        assert err <= 1e-5
        assert h[-1] == 0.0
    if phys == G.TE:
        om2abs = abs(2.0 * np.pi * (freqs[0] + 0j)) ** 2
        hsig, heps = (x.cpu().numpy() for x in hs)
        # Omega and Omega^2 reach the kernel as floats (gfdfd_omega): agreement to their rounding
        assert np.max(np.abs(heps - om2abs * hsig)) <= 1e-5 * np.max(heps)


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
@pytest.mark.parametrize("phys", [G.AC, G.TE])
def test_gauss_newton_hessian_matches_oracle(phys, precision):
    """hessian_gn (src/AC/hessian_AC.c:50-160, src/TE/hessian_TE.c:53-177: every
    shot field against every receiver Green's function) on the device against
    the oracle's per-pair float restatement with FP64 band-LU fields: relative
    1e-5.  TE at one frequency; its epsilon sum exceeds the float range on most
    cells (as in the reference), so it is compared where the float restatement
    is finite and checked everywhere as |Omega|^2 times the sigma sum."""
    say("gauss-newton hessian", phys, precision)
    m = models(phys)
    dh, freqs, n = m["dh"], m["freqs"], NX * NY
    if phys == G.TE:
        freqs = freqs[:1]
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = [O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in (3, 11, 19, 27, 35, 43)]
    pl = fdfd.Plan(NX, NY)
    out = fdfd.hessian_gn(pl, phys, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, recs, npml=NPML,
                          precision=precision, tol=None if precision == G.FP64 else 1e-10)
    hs = [out] if phys == G.AC else list(out)
    refs = [np.zeros(n, np.float32) for _ in hs]
    amp = np.float32(1.0) / np.float32(np.float32(dh) * np.float32(dh))
    for f in freqs:
        _, U = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs)
        rhs = np.zeros((n, len(recs)), complex)
        for r, k in enumerate(recs):
            rhs[k, r] = float(amp)
        _, GR = oracle_fields(phys, dh, f, m["m1"], m["m2"], recs, rhs=rhs)
        for s in range(len(srcs)):
            for r in range(len(recs)):
                if phys == G.AC:
                    O.hess_gn_ac_accumulate(NX, NY, f, 0.0, U[:, s], 0j, GR[:, r], 0j, m["m1"], m["m2"], refs[0])
                else:
                    O.hess_gn_te_accumulate(NX, NY, f, 0.0, U[:, s], 0j, GR[:, r], 0j, refs[0], refs[1])
    for i, (h, ref) in enumerate(zip(hs, refs)):
        h = h.cpu().numpy()
        fin = np.isfinite(ref[:-1])
        assert fin.all() if i == 0 else True
        if fin.any():
            err = float(np.max(np.abs(h[:-1][fin] - ref[:-1][fin])) / np.max(np.abs(ref[:-1][fin])))
            say("  rel err", err, "on", int(fin.sum()), "cells")
            assert err <= 1e-5
        assert h[-1] == 0.0
    if phys == G.TE:
        om2abs = abs(2.0 * np.pi * (freqs[0] + 0j)) ** 2
        hsig, heps = (x.cpu().numpy() for x in hs)
        assert np.max(np.abs(heps - om2abs * hsig)) <= 1e-5 * np.max(heps)


def test_apply_hess_matches_oracle():
    """Diagonal-Hessian preconditioning (src/AC/apply_hess_AC.c) on the device
    against the oracle's float restatement: relative 1e-6."""
    rng = np.random.default_rng(47)
    n = 200001                                     # more cells than the workgroup's threads
    hess = (rng.random(n) ** 4 * 1e3).astype(np.float32).astype(np.float64)
    grad = rng.standard_normal(n).astype(np.float32).astype(np.float64)
    g = dev(grad.copy())
    fdfd.apply_hess(g, dev(hess), 0.005)
    ref = O.apply_hess(0.005, hess.astype(np.float32), grad.astype(np.float32))
    err = float(np.max(np.abs(g.cpu().numpy() - ref) / np.abs(ref).max()))
    say("apply hess rel err", err)
    assert err <= 1e-6


def test_model_update_matches_oracle():
    """Trial model of a line-search step (src/calc_mat_change_wolfe.c:27-46) on
    the device against the oracle: bit-exact, with values crossing both bounds."""
    rng = np.random.default_rng(53)
    n = 100003
    m_old = (1500.0 + 3000.0 * rng.random(n)).astype(np.float32)
    d = (rng.standard_normal(n) * 400.0).astype(np.float64)
    out = fdfd.model_update(1.7, dev(d), dev(m_old), 1600.0, 4400.0)
    ref = O.model_update(1.7, d.astype(np.float32), m_old, 1600.0, 4400.0)
    kept = int(np.count_nonzero(ref == m_old))
    say("model update: cells kept by the bounds", kept)
    assert kept > 1000
    assert np.array_equal(out.cpu().numpy(), ref)


def test_line_search_lowers_the_misfit():
    """One acoustic FWI step search on the device (parabolicls_AC with obj_AC and
    calc_mat_change_wolfe): along the negative gradient the chosen step is
    positive and lowers the misfit, and every trial's misfit equals objective()
    at that trial model."""
    m = models(G.AC)
    dh, freqs = m["dh"], m["freqs"]
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    pl = fdfd.Plan(NX, NY)
    pt = fdfd.forward_shots(pl, G.AC, dh, freqs, dev(m["t1"]), dev(m["t2"]), srcs, recs, npml=NPML, tol=None)
    pobs = pt["data"]
    vp, rho = dev(m["m1"]), dev(m["m2"])
    L2, grad = fdfd.grad_obj(pl, G.AC, dh, freqs, vp, rho, srcs, recs, pobs, npml=NPML, tol=None)
    step, trials = fdfd.line_search(pl, dh, freqs, vp, rho, -grad, srcs, recs, pobs, L2, 0.01, 100.0, 9000.0,
                                    npml=NPML, tol=None)
    say("line search: L2", L2, "trials", trials, "step", step)
    assert step > 0.0
    new = fdfd.model_update(step, -grad, vp, 100.0, 9000.0)
    L2new = fdfd.objective(pl, G.AC, dh, freqs, new, rho, srcs, recs, pobs, npml=NPML, tol=None)
    say("  misfit after the step", L2new)
    assert L2new < L2
    a, l2a = trials[-1]
    again = fdfd.objective(pl, G.AC, dh, freqs, fdfd.model_update(a, -grad, vp, 100.0, 9000.0), rho, srcs, recs,
                           pobs, npml=NPML, tol=None)
    assert abs(again - l2a) <= 1e-12 * abs(l2a)


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_line_search_reuses_factor_storage(precision, monkeypatch):
    """A full acoustic step search (parabolicls_AC, src/AC/parabolicls_AC.c:55-64,
    obj_AC per trial) and the gradient pass of the next iteration on one
    Workspace: one factor arena for every call, each later factorisation a
    replay of the captured refactorisation graph, with eager work and
    synchronisations between replays -- and every trial's misfit equal to a
    fresh factorisation's at the same trial model."""
    say("line search on one workspace", precision)
    m = models(G.AC)
    dh, freqs = m["dh"], m["freqs"]
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    tol = None if precision == G.FP64 else 1e-10
    pl = fdfd.Plan(NX, NY)
    pobs = fdfd.forward_shots(pl, G.AC, dh, freqs, dev(m["t1"]), dev(m["t2"]), srcs, recs, npml=NPML,
                              tol=None)["data"]
    vp, rho = dev(m["m1"]), dev(m["m2"])
    L2, grad = fdfd.grad_obj(pl, G.AC, dh, freqs, vp, rho, srcs, recs, pobs, npml=NPML, precision=precision, tol=tol)
    made = []
    real_init = fdfd.Factors.__init__

    def counting_init(self, *a, **k):
        made.append(1)
        real_init(self, *a, **k)

    monkeypatch.setattr(fdfd.Factors, "__init__", counting_init)
    ws = fdfd.Workspace(pl, len(freqs), precision)
    step, trials = fdfd.line_search(pl, dh, freqs, vp, rho, -grad, srcs, recs, pobs, L2, 0.01, 100.0, 9000.0,
                                    npml=NPML, precision=precision, tol=tol, ws=ws)
    new = fdfd.model_update(step, -grad, vp, 100.0, 9000.0)
    L2b, grad_b = fdfd.grad_obj(pl, G.AC, dh, freqs, new, rho, srcs, recs, pobs, npml=NPML, precision=precision,
                                tol=tol, ws=ws)
    say("  trials", trials, "step", step, "factor arenas", len(made), "misfit after", L2b)
    assert len(made) == 1 and ws.allocations == 1 and len(trials) >= 2
    assert step > 0.0 and L2b < L2
    monkeypatch.setattr(fdfd.Factors, "__init__", real_init)
    # every trial against fresh factor storage at the same trial model
    for a, l2a in trials:
        fresh = fdfd.objective(pl, G.AC, dh, freqs, fdfd.model_update(a, -grad, vp, 100.0, 9000.0), rho, srcs, recs,
                               pobs, npml=NPML, precision=precision, tol=tol)
        assert abs(fresh - l2a) <= 1e-12 * abs(l2a), (a, fresh, l2a)
    L2c, grad_c = fdfd.grad_obj(pl, G.AC, dh, freqs, new, rho, srcs, recs, pobs, npml=NPML, precision=precision,
                                tol=tol)
    assert abs(L2c - L2b) <= 1e-12 * abs(L2b)
    assert float(torch.max(torch.abs(grad_c - grad_b))) <= 1e-12 * float(torch.max(torch.abs(grad_c)))


def test_model_update_norm_matches_oracle():
    """The TE line search's trial model (src/TE/parabolicls_TE.c:30-31,64-69) on
    the device against the oracle: bit-exact, with a normalisation that is not
    a power of two and values clamped at both bounds."""
    rng = np.random.default_rng(59)
    n = 100003
    m_old = (0.002 + 0.002 * rng.random(n)).astype(np.float32)
    d = rng.standard_normal(n) * 0.4
    out = fdfd.model_update_norm(0.9, dev(d), dev(m_old), 0.003, 0.0021, 0.0039)
    ref = O.model_update_norm(0.9, d.astype(np.float32), m_old, 0.003, 0.0021, 0.0039)
    hit = int(np.count_nonzero((ref == ref.min()) | (ref == ref.max())))
    say("model update norm: cells at a bound", hit)
    assert hit > 1000
    assert np.array_equal(out.cpu().numpy(), ref)


def test_line_search_te_lowers_the_misfit():
    """One TE step search on the device (parabolicls_TE without the Tikhonov
    term): along the negative gradients of the normalised materials, each
    scaled to a maximum of one, the step is positive and lowers the misfit."""
    m = models(G.TE)
    dh, freqs = m["dh"], m["freqs"]
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    pl = fdfd.Plan(NX, NY)
    pt = fdfd.forward_shots(pl, G.TE, dh, freqs, dev(m["t1"]), dev(m["t2"]), srcs, recs, npml=NPML, tol=None)
    pobs = pt["data"]
    eps, sig = dev(m["m1"]), dev(m["m2"])
    L2, (g_sig, g_eps) = fdfd.grad_obj(pl, G.TE, dh, freqs, eps, sig, srcs, recs, pobs, npml=NPML, tol=None)
    eps0 = 8.854187817e-12
    norm, low, up = (0.003, 5.0 * eps0), (0.0, eps0), (1.0, 100.0 * eps0)
    d_sig = -g_sig * norm[0]
    d_sig = d_sig / d_sig.abs().max()
    d_eps = -g_eps * norm[1]
    d_eps = d_eps / d_eps.abs().max()
    step, trials = fdfd.line_search_te(pl, dh, freqs, eps, sig, d_eps, d_sig, srcs, recs, pobs, L2, 0.01, norm, low,
                                       up, npml=NPML, tol=None)
    say("TE line search: L2", L2, "trials", trials, "step", step)
    assert step > 0.0
    ns = fdfd.model_update_norm(step, d_sig, sig, norm[0], low[0], up[0])
    ne = fdfd.model_update_norm(step, d_eps, eps, norm[1], low[1], up[1])
    L2new = fdfd.objective(pl, G.TE, dh, freqs, ne, ns, srcs, recs, pobs, npml=NPML, tol=None)
    say("  misfit after the step", L2new)
    assert L2new < L2


@pytest.mark.parametrize("sizes", [(1.0, 0.6, 0), (2.3, 1.7, 6), (0.0, 1.0, 5)])
def test_precond_matches_oracle(sizes):
    """precond's Gaussian filter and horizontal taper (src/precond.c:17-26) on the
    device against the oracle on a random field, including windows that reach
    past every edge and a zero-width window: bit-exact (the weights round exp
    to float on both sides; the sums follow the reference's order in float)."""
    fsx, fsy, nzero = sizes
    ny, nx = 61, 83
    f = np.random.default_rng(61).standard_normal((ny, nx)).astype(np.float32)
    out = fdfd.precond(dev(f.astype(np.float64).ravel()), nx, ny, 10.0, 1500.0, 15.0, fsx, fsy, 1, nzero)
    ref = O.precond(f, 10.0, 1500.0, 15.0, fsx, fsy, 1, nzero).ravel()
    got = out.cpu().numpy()
    exact = float(np.mean(got == ref))
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    say("precond", sizes, "rel err", err, "bit-exact fraction", exact)
    assert np.array_equal(got, ref), err


def test_tikhonov_te_matches_oracle():
    """Tikhonov regularisation of the TE materials on the device against the
    oracle: the gradient terms bit-exact (each is one float product added to
    zero), the costs to 1e-4 (the reference sums them in float, the device in
    double)."""
    rng = np.random.default_rng(67)
    ny, nx = 38, 52
    eps0 = 8.854187817e-12
    sig = (0.002 + 0.002 * rng.random((ny, nx))).astype(np.float32)
    eps = ((5.0 + rng.random((ny, nx))) * eps0).astype(np.float32)
    beta, norm, lam = (0.7, 1.3), (0.003, 5.0 * eps0), (0.01, 0.02)
    gs, ge = torch.zeros(nx * ny, dtype=torch.float64, device="cuda"), torch.zeros(nx * ny, dtype=torch.float64,
                                                                                   device="cuda")
    cs, ce = fdfd.tikhonov_te(dev(sig), dev(eps), 0.04, beta, norm, lam, gs, ge)
    rs, re_ = np.zeros(nx * ny, np.float32), np.zeros(nx * ny, np.float32)
    os_, oe = O.tikhonov_te(sig, eps, 0.04, beta, norm, lam, rs, re_)
    say("tikhonov costs", cs, os_, ce, oe)
    assert abs(cs - os_) <= 1e-4 * abs(os_) and abs(ce - oe) <= 1e-4 * abs(oe)
    assert np.array_equal(gs.cpu().numpy(), rs.astype(np.float64))
    assert np.array_equal(ge.cpu().numpy(), re_.astype(np.float64))


@pytest.mark.parametrize("free_surf", [0, 1])
def test_taper_file_matches_oracle(free_surf):
    """taper_grad_file (src/taper_grad_file.c:22-66) on the device against the
    oracle's loop restatement: bit-exact, with and without a free surface."""
    nx0, ny0, npml = 45, 30, 8
    fs = 0 if free_surf else npml
    nx, ny = nx0 + 2 * npml, ny0 + npml + fs
    rng = np.random.default_rng(71)
    g = rng.standard_normal((ny, nx)).astype(np.float32)
    taper = rng.random((ny0, nx0)).astype(np.float32)
    out = fdfd.taper_file(dev(g.astype(np.float64).ravel()), dev(taper), npml, free_surf)
    ref = O.taper_file(g, taper, npml, free_surf)
    assert np.array_equal(out.cpu().numpy(), ref.ravel().astype(np.float64))


def test_fwi_iteration_ac():
    """Two acoustic FWI iterations in the order of src/AC/fwi_FD_AC.c:263-353 with
    steepest descent (GRAD_METHOD=3) and the pseudo-Hessian (HESSIAN=2): gradient
    and misfit followed by cp_grad_frame (grad_obj_AC, src/AC/grad_obj_AC.c:177),
    the Shin pseudo-Hessian at the first iteration only, also framed
    (src/AC/hessian_shin_AC.c:140), apply_hess_AC every iteration, precond,
    cp_grad_frame and descent (src/AC/fwi_FD_AC.c:287-288), the parabolic step
    search and calc_mat_change_wolfe -- every stage on the device, every
    factorisation in one Workspace.  The misfit must fall at every iteration."""
    m = models(G.AC)
    dh, freqs = m["dh"], m["freqs"]
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    nx0, ny0 = NX - 2 * NPML, NY - 2 * NPML
    pl = fdfd.Plan(NX, NY)
    pobs = fdfd.forward_shots(pl, G.AC, dh, freqs, dev(m["t1"]), dev(m["t2"]), srcs, recs, npml=NPML,
                              tol=None)["data"]
    vp, rho = dev(m["m1"]), dev(m["m2"])
    ws = fdfd.Workspace(pl, len(freqs), G.FP64)
    misfits, hess = [], None
    for it in range(2):
        L2, grad = fdfd.grad_obj(pl, G.AC, dh, freqs, vp, rho, srcs, recs, pobs, npml=NPML, tol=None, ws=ws)
        fdfd.grad_frame(grad, nx0, ny0, NPML, 0)
        if it == 0:
            hess = fdfd.hessian_shin(pl, G.AC, dh, freqs, vp, rho, srcs, npml=NPML, tol=None, ws=ws)
            fdfd.grad_frame(hess, nx0, ny0, NPML, 0)
        fdfd.apply_hess(grad, hess, 0.005)
        grad = fdfd.precond(grad, NX, NY, dh, 1800.0, freqs[0], 0.5, 0.5, 1, NPML)
        fdfd.grad_frame(grad, nx0, ny0, NPML, 0)
        direction = -grad
        step, trials = fdfd.line_search(pl, dh, freqs, vp, rho, direction, srcs, recs, pobs, L2, 0.01, 100.0,
                                        9000.0, npml=NPML, tol=None, ws=ws)
        say("FWI iteration", it, "misfit", L2, "step", step, "trials", len(trials))
        assert step > 0.0
        vp = fdfd.model_update(step, direction, vp, 100.0, 9000.0)
        misfits.append(L2)
    final = fdfd.objective(pl, G.AC, dh, freqs, vp, rho, srcs, recs, pobs, npml=NPML, tol=None, ws=ws)
    say("  misfits", misfits, "final", final, "factor arenas", ws.allocations)
    assert misfits[1] < misfits[0] and final < misfits[1] and ws.allocations == 1


@pytest.mark.parametrize("free_surf", [0, 1])
def test_grad_frame_matches_oracle(free_surf):
    """cp_grad_frame's PML damping (src/cp_grad_frame.c:102-208) on the device
    against the oracle's loop restatement: relative 1e-6 (the npml coefficients
    round a double exp to float on both sides)."""
    nx0, ny0, npml = 45, 30, 8
    fs = 0 if free_surf else npml
    nx, ny = nx0 + 2 * npml, ny0 + npml + fs
    g = np.random.default_rng(73).standard_normal((ny, nx)).astype(np.float32)
    out = fdfd.grad_frame(dev(g.astype(np.float64).ravel()), nx0, ny0, npml, free_surf).cpu().numpy()
    ref = O.grad_frame(g, nx0, ny0, npml, free_surf).ravel()
    say("grad frame", free_surf, "bit-exact fraction", float(np.mean(out == ref)))
    assert np.max(np.abs(out - ref)) <= 1e-6 * np.max(np.abs(ref))


def _te_direction_device(pl, dh, freqs, eps, sig, srcs, recs, pobs, norm, beta, lam, ws):
    """fwi_FD_TE's steps from the gradient to the descent direction
    (src/TE/fwi_FD_TE.c:262-306, GRAD_METHOD=3, HESSIAN=2, LAMBDA > 0) on the device."""
    nx0, ny0 = NX - 2 * NPML, NY - 2 * NPML
    L2, (g_sig, g_eps) = fdfd.grad_obj(pl, G.TE, dh, freqs, eps, sig, srcs, recs, pobs, npml=NPML, tol=None, ws=ws)
    for g, nm in ((g_sig, norm[0]), (g_eps, norm[1])):           # src/TE/grad_obj_TE.c:178-191
        fdfd.grad_frame(g, nx0, ny0, NPML, 0)
        g *= float(np.float32(nm))
    cs, ce = fdfd.tikhonov_te(sig, eps, dh, beta, norm, lam, g_sig, g_eps)
    L2 += lam[0] * cs + lam[1] * ce
    h_sig, h_eps = fdfd.hessian_shin(pl, G.TE, dh, freqs, eps, sig, srcs, npml=NPML, tol=None, ws=ws)
    for h, nm in ((h_sig, norm[0]), (h_eps, norm[1])):           # src/TE/hessian_shin_TE.c:141-148
        fdfd.grad_frame(h, nx0, ny0, NPML, 0)
        h *= float(np.float32(nm) * np.float32(nm))
    fdfd.apply_hess(g_sig, h_sig, 0.005)
    fdfd.apply_hess(g_eps, h_eps, 0.005)
    fdfd.grad_frame(g_sig, nx0, ny0, NPML, 0)
    fdfd.grad_frame(g_eps, nx0, ny0, NPML, 0)
    return L2, -g_sig, -g_eps


def test_te_direction_matches_oracle_sequence():
    """The TE descent direction of one fwi_FD_TE iteration (gradients framed and
    scaled by MAT*_NORM, Tikhonov gradient, Shin pseudo-Hessians framed and
    scaled by NORM^2, apply_hess per material, cp_grad_frame, descent) on the
    device against the oracle's pieces composed in the same order with band-LU
    fields: relative 1e-5 per material.  Low frequencies keep the float
    pseudo-Hessian of the restatement finite; the Tikhonov weights are chosen
    so that the regulariser moves each direction by at least 10%."""
    m = models(G.TE)
    dh, freqs = m["dh"], [1.5e7, 2.5e7]
    n = NX * NY
    nx0, ny0 = NX - 2 * NPML, NY - 2 * NPML
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    eps0 = 8.854187817e-12
    norm, beta = (0.003, 5.0 * eps0), (1.0, 1.0)
    pobs = []
    for f in freqs:
        _, Ut = oracle_fields(G.TE, dh, f, m["t1"], m["t2"], srcs)
        pobs.append(Ut[recs, :].T.astype(np.complex64))
    # oracle: gradients and pseudo-Hessians per shot in float, then the sequence
    gs = np.zeros(n, np.float32)
    ge = np.zeros(n, np.float32)
    hs = np.zeros(n, np.float32)
    he = np.zeros(n, np.float32)
    for fi, f in enumerate(freqs):
        _, U = oracle_fields(G.TE, dh, f, m["m1"], m["m2"], srcs)
        rhs = np.zeros((n, len(srcs)), complex)
        for s_ in range(len(srcs)):
            res, _ = O.residual(U[recs, s_].astype(np.complex64), pobs[fi][s_])
            rhs[:, s_] = O.adjoint_rhs(n, recs, res, dh)
        _, LAM = oracle_fields(G.TE, dh, f, m["m1"], m["m2"], srcs, rhs, 2)
        for s_ in range(len(srcs)):
            O.grad_te_accumulate(NX, NY, f, 0.0, U[:, s_], 0j, LAM[:, s_], 0j, gs, ge)
            O.hess_shin_te_accumulate(NX, NY, f, 0.0, U[:, s_], 0j, hs, he)
    assert np.all(np.isfinite(he)) and np.all(np.isfinite(hs))
    frame = lambda a: O.grad_frame(a.reshape(NY, NX), nx0, ny0, NPML, 0).ravel()
    gs, ge = np.float32(norm[0]) * frame(gs), np.float32(norm[1]) * frame(ge)
    # Tikhonov weights from the regulariser's own gradient (lam = 1)
    ts, te = np.zeros(n, np.float32), np.zeros(n, np.float32)
    O.tikhonov_te(m["m2"], m["m1"], dh, beta, norm, (1.0, 1.0), ts, te)
    lam = (float(0.5 * np.abs(gs).max() / np.abs(ts).max()), float(0.5 * np.abs(ge).max() / np.abs(te).max()))
    g_plain = (gs.copy(), ge.copy())
    O.tikhonov_te(m["m2"], m["m1"], dh, beta, norm, lam, gs, ge)
    hs = np.float32(np.float32(norm[0]) * np.float32(norm[0])) * frame(hs)
    he = np.float32(np.float32(norm[1]) * np.float32(norm[1])) * frame(he)
    dirs_o = []
    for g, h, gp in ((gs, hs, g_plain[0]), (ge, he, g_plain[1])):
        O.apply_hess(0.005, h, g)
        O.apply_hess(0.005, h, gp)
        d, dp = -frame(g), -frame(gp)
        assert np.linalg.norm(d - dp) >= 0.1 * np.linalg.norm(d)     # the regulariser changes the direction
        dirs_o.append(d)
    pl = fdfd.Plan(NX, NY)
    ws = fdfd.Workspace(pl, len(freqs), G.FP64)
    _, d_sig, d_eps = _te_direction_device(pl, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, recs,
                                           [dev(p) for p in pobs], norm, beta, lam, ws)
    for name, dd, do in (("sigma", d_sig, dirs_o[0]), ("eps", d_eps, dirs_o[1])):
        e = float(np.linalg.norm(dd.cpu().numpy() - do) / np.linalg.norm(do))
        say("  TE direction", name, "rel err vs oracle sequence", e)
        assert e <= 1e-5


def test_fwi_iteration_te():
    """Two TE FWI iterations in the order of src/TE/fwi_FD_TE.c:258-345
    (GRAD_METHOD=3, HESSIAN=2, LAMBDA > 0 so precond is skipped), every stage on
    the device and every factorisation in one Workspace: the direction of
    test_te_direction_matches_oracle_sequence, then the normalised step search
    with the Tikhonov term.  The Tikhonov weights make each material's
    Tikhonov cost a tenth of the data misfit at the start model: the
    reference's Tikhonov gradient (LAMBDA BETA lap m, src/TE/Tikhonov_grad_TE.c:
    33-34) is not the gradient of its cost (0.5 BETA sum (lap m)^2,
    src/TE/Tikhonov_cost_TE.c:35-36), so a regulariser that dominates the
    objective cannot be lowered along the reference's direction.  The misfit
    with the Tikhonov term must fall at every iteration."""
    m = models(G.TE)
    dh, freqs = m["dh"], m["freqs"]
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (4, 13, 22, 31, 40)]
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    pl = fdfd.Plan(NX, NY)
    pobs = fdfd.forward_shots(pl, G.TE, dh, freqs, dev(m["t1"]), dev(m["t2"]), srcs, recs, npml=NPML,
                              tol=None)["data"]
    eps, sig = dev(m["m1"]), dev(m["m2"])
    eps0 = 8.854187817e-12
    norm, low, up = (0.003, 5.0 * eps0), (0.0, eps0), (1.0, 100.0 * eps0)
    beta = (1.0, 1.0)
    ws = fdfd.Workspace(pl, len(freqs), G.FP64)
    L2d, _ = fdfd.grad_obj(pl, G.TE, dh, freqs, eps, sig, srcs, recs, pobs, npml=NPML, tol=None, ws=ws)
    cs, ce = fdfd.tikhonov_te(sig, eps, dh, beta, norm, (1.0, 1.0))
    lam = (0.1 * L2d / cs, 0.1 * L2d / ce)
    misfits, final = [], None
    for it in range(2):
        L2, d_sig, d_eps = _te_direction_device(pl, dh, freqs, eps, sig, srcs, recs, pobs, norm, beta, lam, ws)
        # first trial: 1% of max(sigma / MAT1_NORM) at the largest |direction| of either material
        a0 = 0.01 * float(sig.max()) / norm[0] / max(float(d_sig.abs().max()), float(d_eps.abs().max()))
        step, trials = fdfd.line_search_te(pl, dh, freqs, eps, sig, d_eps, d_sig, srcs, recs, pobs, L2, 0.01, norm,
                                           low, up, tikhonov=(beta, lam), npml=NPML, tol=None, ws=ws, alpha0=a0)
        say("TE FWI iteration", it, "misfit", L2, "step", step, "trials", len(trials))
        assert step > 0.0
        sig = fdfd.model_update_norm(step, d_sig, sig, norm[0], low[0], up[0])
        eps = fdfd.model_update_norm(step, d_eps, eps, norm[1], low[1], up[1])
        misfits.append(L2)
    cs, ce = fdfd.tikhonov_te(sig, eps, dh, beta, norm, lam)
    final = fdfd.objective(pl, G.TE, dh, freqs, eps, sig, srcs, recs, pobs, npml=NPML, tol=None, ws=ws)
    final += lam[0] * cs + lam[1] * ce
    say("  misfits", misfits, "final (with the Tikhonov term)", final, "lam", lam)
    assert misfits[1] < misfits[0] and final < misfits[1] and ws.allocations == 1


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("phys", [G.AC, G.TE])
def test_gradient_shot_taper_matches_oracle(phys, mode):
    """ass_grad's per-shot taper (SWS_TAPER_CIRCULAR_PER_SHOT, src/AC/ass_grad_AC.c:37-43,
    src/TE/ass_grad_TE.c:40-49): the oracle tapers every shot's float gradient with
    taper_grad_shot (mode 1: SRTRADIUS 30 m around the shot's source and every receiver) or
    taper_grad_shot1 (mode 2: FILTSIZE 2 boxes around every source and receiver) before
    adding it, in the reference's coordinates (cell (ix, iy) at (ix DH, iy DH) against
    srcpos in metres and recpos in grid indices); the device applies the same factors
    inside its shot sum."""
    say("gradient shot taper", phys, mode)
    m = models(phys)
    dh, freqs, n = m["dh"], m["freqs"], NX * NY
    xs = (4, 13, 22, 31, 40)
    srcpos = np.array([[x * dh, 12 * dh] for x in xs], np.float32)
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in xs]
    recpos = np.array([[x, 10] for x in range(1, 44)], np.int32)
    recs = np.array([O.source_index(x * dh, 10 * dh, dh, NPML, 0, NX) for x in range(1, 44)], np.int32)
    pobs = []
    for f in freqs:
        _, Ut = oracle_fields(phys, dh, f, m["t1"], m["t2"], srcs)
        pobs.append(Ut[recs, :].T.astype(np.complex64))
    radius, filtsize = 3.0 * dh, 2
    taper = fdfd.ShotTaper(mode, NX, dh, srcpos, recpos, srtradius=radius, filtsize=filtsize)
    pl = fdfd.Plan(NX, NY)
    _, grad = fdfd.grad_obj(pl, phys, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, recs, [dev(p) for p in pobs],
                            npml=NPML, precision=G.FP64, tol=None, taper=taper)
    torch.cuda.synchronize()
    g1 = np.zeros(n, np.float32)
    g2 = np.zeros(n, np.float32)
    for fi, f in enumerate(freqs):
        csc, U = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs)
        rhs = np.zeros((n, len(srcs)), complex)
        for s in range(len(srcs)):
            res, _ = O.residual(U[recs, s].astype(np.complex64), pobs[fi][s])
            rhs[:, s] = O.adjoint_rhs(n, recs, res, dh)
        _, LAM = oracle_fields(phys, dh, f, m["m1"], m["m2"], srcs, rhs, m["adj"])
        for s in range(len(srcs)):
            shot = [np.zeros(n, np.float32), np.zeros(n, np.float32)]
            if phys == G.AC:
                O.grad_ac_accumulate(NX, NY, f, 0.0, U[:, s], 0j, LAM[:, s], 0j, m["m1"], m["m2"], shot[0])
            else:
                O.grad_te_accumulate(NX, NY, f, 0.0, U[:, s], 0j, LAM[:, s], 0j, shot[0], shot[1])
            for g in shot:
                if mode == 1:
                    O.taper_grad_shot(g, NX, NY, dh, radius, srcpos[s, 0], srcpos[s, 1], recpos[:, 0], recpos[:, 1])
                else:
                    O.taper_grad_shot1(g, NX, NY, dh, filtsize, srcpos[:, 0], srcpos[:, 1], recpos[:, 0],
                                       recpos[:, 1])
            g1 += shot[0]
            g2 += shot[1]
    got = [grad.cpu().numpy()] if phys == G.AC else [grad[0].cpu().numpy(), grad[1].cpu().numpy()]
    plain = fdfd.grad_obj(pl, phys, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, recs, [dev(p) for p in pobs],
                          npml=NPML, precision=G.FP64, tol=None)[1]
    plain = [plain.cpu().numpy()] if phys == G.AC else [plain[0].cpu().numpy(), plain[1].cpu().numpy()]
    for d, r, u in zip(got, [g1, g2], plain):
        e = float(np.linalg.norm(d - r) / np.linalg.norm(r))
        changed = int(np.sum(np.abs(u - d) > 1e-3 * np.abs(u)))
        say("  gradient rel err", e, "cells the taper changed", changed)
        assert e <= 1e-5
        assert changed > 0
        if mode == 2:
            # a box cell holds 1e-5 per shot and frequency, summed in float by the oracle
            boxed = np.abs(r - 1e-5 * len(freqs) * len(srcs)) <= 1e-9
            assert boxed.sum() >= len(srcs) * 25
            assert np.allclose(d[boxed], r[boxed], rtol=1e-6, atol=0)


def test_wavelet_source_matches_oracle():
    """READ_WAVELET = 1 (src/AC/RHS_source_AC.c:41-43): the forward panel with the
    wavelet's DFT as source amplitude, against the oracle's band solve with the
    oracle's DFT of the same trace."""
    say("wavelet source")
    m = models(G.AC)
    dh, freqs = m["dh"], m["freqs"]
    t = np.arange(400, dtype=np.float32) * np.float32(1e-3)
    w = ((1 - 2 * (np.pi * 8.0 * (t - 0.1)) ** 2) * np.exp(-(np.pi * 8.0 * (t - 0.1)) ** 2)).astype(np.float32)
    srcs = [O.source_index(x * dh, 12 * dh, dh, NPML, 0, NX) for x in (6, 25, 44)]
    pl = fdfd.Plan(NX, NY)
    out = fdfd.forward_shots(pl, G.AC, dh, freqs, dev(m["m1"]), dev(m["m2"]), srcs, npml=NPML, precision=G.FP64,
                             tol=None, keep_fields=True, wavelet=(w, 1000))
    for fi, f in enumerate(freqs):
        amp = O.wavelet_amp(G.AC, f, dh, w, 1000)
        assert amp == fdfd.source_amp(G.AC, f, 0.0, dh, (w, 1000))
        rhs = np.zeros((NX * NY, len(srcs)), complex)
        for s, k in enumerate(srcs):
            rhs[k, s] = amp
        _, X = oracle_fields(G.AC, dh, f, m["m1"], m["m2"], srcs, rhs)
        got = out["fields"][fi].cpu().numpy().T
        e = float(np.max(np.abs(got - X)) / np.max(np.abs(X)))
        say("  freq", f, "amp", amp, "rel err", e)
        assert e <= 1e-10
