"""Full-size parity of the benchmark workloads on the device.

For each workload at its real grid, frequencies and shot count:

* every shot's wavefield is checked by its FP64 relative residual
  ||b - M x|| / ||b||, computed by the nine-diagonal operator of
  tests/_stencil.py (elementwise products, independent of the library's
  residual kernels), built from the oracle's assembly;
* sampled shots are checked against the oracle's FP64 multifrontal LU
  (oracle/mf_cpu.c, OpenMP on the host cores): relative wavefield error
  <= 1e-10 for FP64 factors, <= 1e-8 for FP32 factors with FP64 refinement;
* the FWI workload's misfit and gradient against the oracle's per-shot
  restatement of grad_obj_AC (src/AC/grad_obj_AC.c:50-168) with mf_cpu
  solves, relative 1e-5, then five l-BFGS iterations on one Workspace.

Workloads (grids include the PML): ac_801x187_4f, ac_1601x401_8f,
ac_801x201_fwi, te_2401x801_6f.  Each test prints REPORT lines (JSON) with the
per-frequency errors, refinement steps and, for the TE workload, the device
memory the library held at its peak.
"""
import json
import os
import sys
import time

import numpy as np
import pytest

import germaine_amd as G
from germaine_amd import fdfd, models
from oracle import mf
from oracle import oracle as O
from tests import _stencil as ST

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def report(name, **kv):
    line = json.dumps({"workload": name, **kv})
    say("REPORT", line)
    d = os.environ.get("GFDFD_REPORT_DIR")
    if d:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "config_reports.jsonl"), "a") as f:
            f.write(line + "\n")


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def oracle_triplets(w, f):
    px, py = O.pml_profiles(w.nx, w.ny, w.npml, w.dh, w.a0_pml, f, w.s, 0)
    return O.assemble(w.physics, w.nx, w.ny, w.dh, f, w.s, w.m1, w.m2, px, py)


def spike_rhs(w, f, shots, n):
    amp = O.source_amp(w.physics, f, w.s, w.dh)
    B = np.zeros((len(shots), n), complex)
    for r, k in enumerate(shots):
        B[r, k] = amp
    return B


def device_relres(tri, n, X, ksrc, amp):
    """max over the shots of X [nshots, n] (cuda) of ||b - M x|| / ||b||."""
    D = ST.to_device(ST.diagonals(n, *tri), torch)
    out = []
    for c0 in range(0, X.shape[0], 50):
        Xc = X[c0:c0 + 50]
        B = torch.zeros_like(Xc)
        rows = torch.arange(Xc.shape[0], device="cuda")
        B[rows, torch.as_tensor(np.asarray(ksrc[c0:c0 + 50]), device="cuda")] = amp
        out.append(ST.relres(D, Xc, B))
    return float(torch.cat(out).max())


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sample_idx(count, k):
    return sorted(set(int(round(i * (count - 1) / (k - 1))) for i in range(k)))


def workload_tensors(w):
    return dev(w.m1), dev(w.m2)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_ac_801x187_4f(precision):
    """801x187 acoustic, 20-cell PML, h = 25 m, layered-fault model (seed 2), 3/5/7/10 Hz,
    100 shots and 400 receivers at the surface: every shot's relative residual, and 5 sampled
    shots per frequency (wavefield and receiver data) against the FP64 oracle."""
    w = models.config("ac_801x187_4f")
    n = w.nx * w.ny
    say("workload", w.name, "precision", precision)
    m1, m2 = workload_tensors(w)
    pl = fdfd.Plan(w.nx, w.ny)
    t0 = time.perf_counter()
    out = fdfd.forward_shots(pl, w.physics, w.dh, w.freqs, m1, m2, w.src, w.rec, npml=w.npml, precision=precision,
                             tol=None if precision == G.FP64 else 1e-10, keep_fields=True)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    mf.set_threads(min(16, os.cpu_count() or 1))
    per_freq = []
    for fi, f in enumerate(w.freqs):
        tri = oracle_triplets(w, f)
        amp = O.source_amp(w.physics, f, w.s, w.dh)
        rr = device_relres(tri, n, out["fields"][fi], w.src, amp)
        F = mf.Factor(w.nx, w.ny, *O.csc_of_M(w.nx, w.ny, *tri))
        assert not F.singular
        idx = sample_idx(len(w.src), 5)
        Xo = F.solve(spike_rhs(w, f, [w.src[i] for i in idx], n), 0)
        Xd = out["fields"][fi][idx].cpu().numpy()
        err = max(rel(Xd[r], Xo[r]) for r in range(len(idx)))
        data = out["data"][fi][idx].cpu().numpy()
        derr = max(rel(data[r], Xo[r][w.rec].astype(np.complex64)) for r in range(len(idx)))
        per_freq.append(dict(freq=f, max_relres_all_shots=rr, max_rel_err_vs_oracle=err, data_rel_err=derr,
                             ir_steps=out["iters"][fi]))
        say(f"  {f} Hz: relres {rr:.2e}, wavefield err vs oracle {err:.2e}, receivers {derr:.2e}, "
            f"IR steps {out['iters'][fi]}")
        assert rr <= (1e-12 if precision == G.FP64 else 1e-10)
        assert err <= (1e-10 if precision == G.FP64 else 1e-8)
        assert derr <= 1e-6
    report(w.name, precision="fp64" if precision == G.FP64 else "fp32+ir", shots=len(w.src), receivers=len(w.rec),
           device_s=round(t_dev, 3), per_freq=per_freq)


@pytest.mark.timeout(900)
def test_ac_1601x401_8f_batched():
    """1601x401 acoustic, 30-cell PML, h = 12.5 m, 8 frequencies 2-16 Hz factored in one batch,
    200-shot panels: FP32 factors with FP64 refinement to a relative residual of 1e-10 and FP64
    factors, every shot's relative residual for both, and 6 sampled shots at 2 and 16 Hz against
    the FP64 oracle."""
    w = models.config("ac_1601x401_8f")
    n, nf, ns = w.nx * w.ny, len(w.freqs), len(w.src)
    say("workload", w.name)
    m1, m2 = workload_tensors(w)
    pl = fdfd.Plan(w.nx, w.ny)
    vals = pl.assemble(w.physics, w.dh, w.freqs, m1, m2, npml=w.npml)
    B = torch.stack([fdfd.spike_panel(n, w.src, fdfd.source_amp(w.physics, f, w.s, w.dh)) for f in w.freqs])
    res = {}
    for prec in (G.FP32, G.FP64):
        F = pl.factor(vals, prec)
        if prec == G.FP32:
            X, its, rrs = F.solve_refine_batch(list(range(nf)), B, tol=1e-10, max_iter=10)
        else:
            X, its, rrs = F.solve_batch(list(range(nf)), B), [0] * nf, [float("nan")] * nf
        torch.cuda.synchronize()
        del F
        res[prec] = (X, its, rrs)
    mf.set_threads(min(16, os.cpu_count() or 1))
    per_freq = []
    idx = sample_idx(ns, 6)
    for fi, f in enumerate(w.freqs):
        tri = oracle_triplets(w, f)
        amp = O.source_amp(w.physics, f, w.s, w.dh)
        row = dict(freq=f, ir_steps=res[G.FP32][1][fi])
        for prec, name in ((G.FP32, "fp32+ir"), (G.FP64, "fp64")):
            row[f"max_relres_{name}"] = device_relres(tri, n, res[prec][0][fi], w.src, amp)
        row["fp32ir_vs_fp64"] = float(torch.max(torch.abs(res[G.FP32][0][fi] - res[G.FP64][0][fi])) /
                                      torch.max(torch.abs(res[G.FP64][0][fi])))
        if f in (w.freqs[0], w.freqs[-1]):
            F = mf.Factor(w.nx, w.ny, *O.csc_of_M(w.nx, w.ny, *tri))
            Xo = F.solve(spike_rhs(w, f, [w.src[i] for i in idx], n), 0)
            for prec, name in ((G.FP32, "fp32+ir"), (G.FP64, "fp64")):
                Xd = res[prec][0][fi][idx].cpu().numpy()
                row[f"max_rel_err_vs_oracle_{name}"] = max(rel(Xd[r], Xo[r]) for r in range(len(idx)))
            del F
        say("  ", row)
        per_freq.append(row)
        assert row["max_relres_fp32+ir"] <= 1e-10 and row["max_relres_fp64"] <= 1e-11
        assert row.get("max_rel_err_vs_oracle_fp32+ir", 0) <= 1e-8
        assert row.get("max_rel_err_vs_oracle_fp64", 0) <= 1e-10
    report(w.name, shots=ns, per_freq=per_freq)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_ac_801x201_fwi_gradient_and_lbfgs(precision):
    """801x201 acoustic FWI workload (h = 20 m, 20-cell PML, 3/5/7/9 Hz, 80 shots, 320 receivers):
    observed data from an FP64 forward run on the true layered-fault model (seed 4), start model
    the true one smoothed by a Gaussian of 15 cells; misfit and gradient at the start model against
    the oracle's per-shot restatement of grad_obj_AC with FP64 multifrontal solves (relative 1e-5;
    the reference accumulates in float), then five l-BFGS iterations (memory 5) with the parabolic
    step search, refactoring every frequency for every evaluation on one Workspace: the misfit
    must fall at every iteration."""
    w = models.config("ac_801x201_fwi")
    n = w.nx * w.ny
    say("workload", w.name, "precision", precision)
    vp_true, rho = dev(w.m1), dev(w.m2)
    vp0_np = models.gaussian_smooth(w.m1, 15.0)
    vp0 = dev(vp0_np)
    pl = fdfd.Plan(w.nx, w.ny)
    pobs = fdfd.forward_shots(pl, w.physics, w.dh, w.freqs, vp_true, rho, w.src, w.rec, npml=w.npml,
                              tol=None)["data"]
    tol = None if precision == G.FP64 else 1e-10
    ws = fdfd.Workspace(pl, len(w.freqs), precision)
    t0 = time.perf_counter()
    L2, grad = fdfd.grad_obj(pl, w.physics, w.dh, w.freqs, vp0, rho, w.src, w.rec, pobs, npml=w.npml,
                             precision=precision, tol=tol, ws=ws)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    # oracle: per frequency one FP64 multifrontal factorisation, forward and adjoint solves per shot
    mf.set_threads(min(16, os.cpu_count() or 1))
    g = np.zeros(n, np.float32)
    l2 = 0.0
    recs = np.asarray(w.rec, np.int32)
    for fi, f in enumerate(w.freqs):
        px, py = O.pml_profiles(w.nx, w.ny, w.npml, w.dh, w.a0_pml, f, w.s, 0)
        tri = O.assemble(w.physics, w.nx, w.ny, w.dh, f, w.s, vp0_np, w.m2, px, py)
        F = mf.Factor(w.nx, w.ny, *O.csc_of_M(w.nx, w.ny, *tri))
        U = F.solve(spike_rhs(w, f, w.src, n), 0)
        P = pobs[fi].cpu().numpy()
        R = np.zeros((len(w.src), n), complex)
        for s_ in range(len(w.src)):
            res_, l2s = O.residual(U[s_, recs].astype(np.complex64), P[s_])
            l2 += l2s
            R[s_] = O.adjoint_rhs(n, recs, res_, w.dh)
        LAM = F.solve(R, 0)
        for s_ in range(len(w.src)):
            O.grad_ac_accumulate(w.nx, w.ny, f, w.s, U[s_], 0j, LAM[s_], 0j, vp0_np, w.m2, g)
        del F
    l2 *= 0.5
    e_grad = rel(grad.cpu().numpy(), g.astype(np.float64))
    e_l2 = abs(L2 - l2) / abs(l2)
    say(f"  misfit {L2:.6e} (oracle {l2:.6e}, rel {e_l2:.1e}), gradient rel err {e_grad:.2e}, "
        f"device gradient pass {t_grad:.3f} s")
    assert e_l2 <= 1e-5 and e_grad <= 1e-5
    # five l-BFGS iterations in fwi_FD_AC's order (src/AC/fwi_FD_AC.c:263-353, HESSIAN=0)
    nx0, ny0 = w.nx - 2 * w.npml, w.ny - 2 * w.npml
    lb = fdfd.LBFGS(memory=5)
    vp = vp0
    misfits = []
    t0 = time.perf_counter()
    for it in range(5):
        if it > 0:
            L2, grad = fdfd.grad_obj(pl, w.physics, w.dh, w.freqs, vp, rho, w.src, w.rec, pobs, npml=w.npml,
                                     precision=precision, tol=tol, ws=ws)
        fdfd.grad_frame(grad, nx0, ny0, w.npml, 0)
        # smoothing over half a wavelength and the rows down to the sources zeroed (taper_grad_hor)
        grad = fdfd.precond(grad, w.nx, w.ny, w.dh, 4000.0, w.freqs[0], 0.5, 0.5, 1, w.npml + 3)
        fdfd.grad_frame(grad, nx0, ny0, w.npml, 0)
        direction = lb.direction(grad, vp)
        fdfd.grad_frame(direction, nx0, ny0, w.npml, 0)
        # first trial: a change of 1% of max(vp) at the largest |direction| (the reference's
        # signed-maximum rule gives no usable step for a direction that is almost all negative)
        a0 = 0.01 * float(vp.max()) / float(direction.abs().max())
        step, trials = fdfd.line_search(pl, w.dh, w.freqs, vp, rho, direction, w.src, w.rec, pobs, L2, 0.01,
                                        1500.0, 7000.0, npml=w.npml, precision=precision, tol=tol, ws=ws,
                                        alpha0=a0)
        say(f"  l-BFGS iteration {it + 1}: misfit {L2:.6e}, step {step:.4g}, {len(trials)} trials")
        assert step > 0.0
        vp = fdfd.model_update(step, direction, vp, 1500.0, 7000.0)
        misfits.append(L2)
    final = fdfd.objective(pl, w.physics, w.dh, w.freqs, vp, rho, w.src, w.rec, pobs, npml=w.npml,
                           precision=precision, tol=tol, ws=ws)
    torch.cuda.synchronize()
    t_fwi = time.perf_counter() - t0
    misfits.append(final)
    say("  misfits", misfits, "factor arenas", ws.allocations, f"{t_fwi:.2f} s")
    assert all(b < a for a, b in zip(misfits, misfits[1:])) and ws.allocations == 1
    mdiff0 = float(torch.linalg.norm(vp0 - vp_true))
    mdiff = float(torch.linalg.norm(vp - vp_true))
    report(w.name, precision="fp64" if precision == G.FP64 else "fp32+ir", shots=len(w.src),
           receivers=len(w.rec), misfit_rel_err=e_l2, gradient_rel_err=e_grad, gradient_pass_s=round(t_grad, 3),
           lbfgs_misfits=misfits, model_error_start=mdiff0, model_error_end=mdiff, fwi_5_iterations_s=round(t_fwi, 2))


@pytest.mark.timeout(1200)
def test_te_2401x801_6f_batched():
    """TE georadar workload 2401x801 (h = 0.04 m, 40-cell PML), relative permittivity 1-25 in
    ellipses over layers, conductivity 0-0.05 S/m, 6 frequencies 100-600 MHz factored in one
    batch, a 150-shot panel per frequency: FP32 factors with FP64 refinement (per-frequency
    refinement steps) and FP64 factors; every shot's relative residual for both; 4 sampled shots
    at 100 and 600 MHz against the FP64 oracle; the device memory the library held at its peak
    (factors, plan and solve buffers) beside torch's (values, right-hand sides, solutions)."""
    w = models.config("te_2401x801_6f")
    n, nf, ns = w.nx * w.ny, len(w.freqs), len(w.src)
    say("workload", w.name)
    m1, m2 = workload_tensors(w)
    pl = fdfd.Plan(w.nx, w.ny)
    vals = pl.assemble(w.physics, w.dh, w.freqs, m1, m2, npml=w.npml)
    amps = [fdfd.source_amp(w.physics, f, w.s, w.dh) for f in w.freqs]
    tris = [oracle_triplets(w, f) for f in w.freqs]
    rows = [dict(freq=f) for f in w.freqs]
    X32 = []
    peaks, times = {}, {}
    cur, peak = C_ll(), C_ll()
    idx = sample_idx(ns, 4)
    for prec, name in ((G.FP32, "fp32+ir"), (G.FP64, "fp64")):
        # solve buffers cached by earlier workloads in this process are released first
        G.check(G.lib().gfdfd_release_workspace(), "gfdfd_release_workspace")
        G.lib().gfdfd_device_memory(None, None, 1)
        torch.cuda.reset_peak_memory_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        F = pl.factor(vals, prec)
        torch.cuda.synchronize()
        times[name + "_factor_6f"] = round(time.perf_counter() - t0, 3)
        t_sol = 0.0
        for fi in range(nf):
            B = fdfd.spike_panel(n, w.src, amps[fi])
            t0 = time.perf_counter()
            if prec == G.FP32:
                X, its, rrs = F.solve_refine_batch([fi], B[None], tol=1e-10, max_iter=10)
                X = X[0]
                rows[fi]["ir_steps"], rows[fi]["ir_relres"] = its[0], rrs[0]
            else:
                X = F.solve(fi, B)
            torch.cuda.synchronize()
            t_sol += time.perf_counter() - t0
            rows[fi][f"max_relres_{name}"] = device_relres(tris[fi], n, X, w.src, amps[fi])
            rows[fi][f"sampled_{name}"] = X[idx].cpu().numpy()
            if prec == G.FP32:
                X32.append(X)
            else:
                rows[fi]["fp32ir_vs_fp64"] = float(torch.max(torch.abs(X32[fi] - X)) / torch.max(torch.abs(X)))
                X32[fi] = None
            del X, B
        times[name + "_solves_6f"] = round(t_sol, 3)
        G.lib().gfdfd_device_memory(cur, peak, 0)
        peaks[name] = dict(library_gb=round(peak.value / 1e9, 2),
                           torch_gb=round(torch.cuda.max_memory_allocated() / 1e9, 2))
        del F
        say("  ", name, times, "peak", peaks[name])
    mf.set_threads(min(16, os.cpu_count() or 1))
    for fi, f in enumerate(w.freqs):
        row = rows[fi]
        if f in (w.freqs[0], w.freqs[-1]):
            F = mf.Factor(w.nx, w.ny, *O.csc_of_M(w.nx, w.ny, *tris[fi]))
            Xo = F.solve(spike_rhs(w, f, [w.src[i] for i in idx], n), 0)
            for name in ("fp32+ir", "fp64"):
                row[f"max_rel_err_vs_oracle_{name}"] = max(rel(row[f"sampled_{name}"][r], Xo[r])
                                                          for r in range(len(idx)))
            del F
        for name in ("fp32+ir", "fp64"):
            del row[f"sampled_{name}"]
        say("  ", row)
    report(w.name, shots=ns, seconds=times, peak_device_memory=peaks, per_freq=rows)
    for row in rows:
        assert row["max_relres_fp32+ir"] <= 1e-10
        assert row["max_relres_fp64"] <= 1e-9
        assert row.get("max_rel_err_vs_oracle_fp32+ir", 0) <= 1e-8
        assert row.get("max_rel_err_vs_oracle_fp64", 0) <= 1e-10


def C_ll():
    import ctypes
    return ctypes.c_longlong(0)
