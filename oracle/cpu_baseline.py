"""ORACLE — test infrastructure only: the CPU baseline leg of bench.py.

Times the CPU restatement of the path on the host cores, on a bounded sample
of the benchmark workload: one frequency (the highest, whose PML profiles
and stencil are the same work as any other) is assembled by the oracle
assembly (oracle/assemble.c, one thread), factorised by the OpenMP FP64
multifrontal LU (oracle/mf_cpu.c) and solved for `sample_shots` point
sources one right-hand side at a time, shots spread over the threads the way
the reference spreads them over MPI ranks (src/AC/forward_AC.c:134-147).
The step time of the full workload is extrapolated as
    nfreq * (t_assemble + t_factor + nshots * t_solve_per_shot)
and reported as shot-solves/s, the benchmark's unit.  One warm-up repetition
of factor + solves, then the median of `reps`.  The sampled shots' FP64
solutions are returned too, so the caller can measure its own solutions'
error against them.
"""
import os
import time

import numpy as np

from oracle import mf
from oracle import oracle as O


def measure(w, threads=None, sample_shots=None, reps=5):
    """Returns (baseline dict, sampled shot indices, their FP64 solutions [k, n])."""
    threads = mf.set_threads(threads or min(16, os.cpu_count() or 1))
    sample_shots = sample_shots or 2 * threads
    f = w.freqs[-1]
    t0 = time.perf_counter()
    px, py = O.pml_profiles(w.nx, w.ny, w.npml, w.dh, w.a0_pml, f, w.s, 0)
    ap, ai, ax, az = O.csc_of_M(w.nx, w.ny, *O.assemble(w.physics, w.nx, w.ny, w.dh, f, w.s, w.m1, w.m2, px, py))
    t_asm = time.perf_counter() - t0
    n = w.nx * w.ny
    idx = [i * len(w.src) // sample_shots for i in range(sample_shots)]
    amp = O.source_amp(w.physics, f, w.s, w.dh)
    B = np.zeros((sample_shots, n), complex)
    for r, i in enumerate(idx):
        B[r, w.src[i]] = amp      # point source at grid row k (src/AC/RHS_source_AC.c:23-32)
    t_facs, t_sols = [], []
    for rep in range(reps + 1):   # the first repetition is the warm-up
        t0 = time.perf_counter()
        F = mf.Factor(w.nx, w.ny, ap, ai, ax, az)
        t_fac = time.perf_counter() - t0
        t0 = time.perf_counter()
        X = F.solve(B, 0)
        t_sol = (time.perf_counter() - t0) / sample_shots
        if rep:
            t_facs.append(t_fac)
            t_sols.append(t_sol)
        flops = F.flops
        del F
    t_fac, t_sol = float(np.median(t_facs)), float(np.median(t_sols))
    rel = float(np.linalg.norm(O.matvec(ap, ai, ax, az, X[0], 0) - B[0]) / np.linalg.norm(B[0]))
    nf, ns = len(w.freqs), len(w.src)
    t_step = nf * (t_asm + t_fac + ns * t_sol)
    base = {"value": round(nf * ns / t_step, 3), "unit": "shot-solves/s", "cores": threads, "kind": "port",
            "sample": f"{w.name} at {f:g} Hz: the oracle's assembly and FP64 OpenMP multifrontal LU (UMFPACK is "
                      f"absent on the box) -- factor + {sample_shots} single-RHS solves, 1 warm-up + median of "
                      f"{reps}, extrapolated to {nf} freqs x {ns} shots",
            "t_assemble_s": round(t_asm, 3), "t_factor_s": round(t_fac, 3), "t_solve_per_shot_s": round(t_sol, 5),
            "factor_gflop": round(flops / 1e9, 2), "relres_shot0": rel}
    return base, idx, X
