#!/usr/bin/env python3
"""Benchmark: shot-solves/s of the FDFD solve path on one MI355X.

Workload (default): 1601 x 401 acoustic grid incl. a 30-cell PML, h = 12.5 m,
seeded layered-fault model, 8 frequencies 2..16 Hz, 200 point sources.
A step is one pass of the hot path over the batch, timed from "model in HBM"
to "all solution panels in HBM":
  device assembly of the 8 matrices -> one batched factorisation of the 8
  frequencies -> per frequency one 200-column point-source panel and its
  multi-RHS solve with FP64 iterative refinement to ||r||/||b|| <= 1e-10.
value = 8 * 200 / (seconds per step), FP32 fronts (FP32 MFMA) + FP64 refinement.

Also reported: the same step with FP64 fronts (no refinement), the plain FP64
path the drop-in ABI executes (per frequency a factorisation and one
single-RHS solve per shot, device-resident), the FP32+IR wavefield error
against the FP64 solution, per-kernel event timings with algorithmic
flops/bytes (every class over the warm-up step; the dominant class over the
timed steps, which gives the roofline), and the CPU baseline (the oracle on
the host cores, on a bounded sample).  Kernel selection comes from the
autotuner (germaine_amd/autotune.py): its cached choice for this workload and
device, or a search over the candidates before anything is timed.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64_mfma": 78.6e12, "fp32_mfma": 157.3e12, "hbm": 8.0e12}
KERNELS = ["assemble", "front_assemble", "scatter_orig", "small_lu", "panel", "panel_apply", "trsm_u", "gemm_update", "diag_inv",
           "fwd_gather", "fwd_leaf", "fwd_small1", "fwd_front1", "fwd_front", "fwd_blk", "bwd_small1", "bwd_front1", "bwd_front",
           "bwd_contrib", "bwd_blk",
           "piv_unswap", "fwd_vec", "bwd_vec", "permute_in", "permute_out", "residual", "norm_reduce", "axpy", "subtree"]
# kernels whose arithmetic runs on the matrix cores (at the factor precision)
MFMA_KERNELS = {"small_lu", "subtree", "panel", "trsm_u", "gemm_update", "fwd_leaf", "fwd_small1", "fwd_front1", "fwd_front", "fwd_blk",
                "bwd_small1", "bwd_front1", "bwd_front", "bwd_contrib", "bwd_blk"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_stats(lib, ctypes):
    out = {}
    for k in KERNELS:
        ms, n, fl, by = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
        lib.gfdfd_profile_read(k.encode(), ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl), ctypes.byref(by))
        if n.value:
            out[k] = dict(ms=ms.value, launches=n.value, flops=fl.value, bytes=by.value)
    return out


def roofline(name, d, mfma_peak):
    """Roofline position of one kernel class from its algorithmic flops and
    bytes per launch and its average event-timed launch duration: the bound is
    whichever of the MFMA and HBM times is larger at peak."""
    avg_s = d["ms"] / d["launches"] * 1e-3
    fl, by = d["flops"] / d["launches"], d["bytes"] / d["launches"]
    peak_f = mfma_peak if name in MFMA_KERNELS else PEAK["fp64_mfma"]
    if fl / peak_f > by / PEAK["hbm"]:
        return dict(bound="mfma", achieved=fl / avg_s / 1e12, peak=peak_f / 1e12, unit="TFLOP/s",
                    frac=fl / avg_s / peak_f, traffic=None, kernel=name, avg_launch_us=avg_s * 1e6,
                    flops_per_launch=fl, bytes_per_launch=by)
    return dict(bound="hbm", achieved=by / avg_s / 1e9, peak=PEAK["hbm"] / 1e9, unit="GB/s",
                frac=by / avg_s / PEAK["hbm"], traffic=None, kernel=name, avg_launch_us=avg_s * 1e6,
                flops_per_launch=fl, bytes_per_launch=by)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(kernel, workload):
    """Copy-calibrated FETCH_SIZE + WRITE_SIZE bytes per launch of a kernel class,
    recorded from rocprofv3 pmc runs of this bench on the same workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("workload_name") != workload:
        return None
    return dict(rec, source=os.path.relpath(TRAFFIC_FILE, ROOT))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="ac_1601x401_8f")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-plain", action="store_true", help="skip the per-shot plain FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 comparison")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-fp64", action="store_true", help="skip the batched FP64 comparison step")
    ap.add_argument("--tune", choices=["auto", "search", "off"], default="auto",
                    help="kernel selection: auto = the autotuner's cached choice for this workload and device, "
                         "searching (before the timed region) when there is none; search = always search; "
                         "off = the switches below")
    ap.add_argument("--small-max", type=int, default=64, help="largest front order factored wholly in LDS")
    ap.add_argument("--fused-rw", type=int, default=64, help="widest fused solve tile (16, 32, 64)")
    ap.add_argument("--fused-mmax", type=int, default=0,
                    help="largest front order of a multi-block level on the fused solve kernels "
                         "(0: none, -1: as far as LDS allows)")
    ap.add_argument("--refine-layout", type=int, default=0,
                    help="panel refinement layout: 0 elimination order (default), 1 caller's layout")
    ap.add_argument("--persistent", type=int, default=1,
                    help="small subtrees factored in one persistent launch (1) or level by level (0)")
    ap.add_argument("--deep-min-m", type=int, default=0,
                    help="blocked levels with a front of order >= this update the trailing block once per 64 "
                         "pivots (0: every level; 1073741824: none)")
    ap.add_argument("--no-graph", action="store_true",
                    help="refactor with eager launches instead of replaying the captured hipGraph")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the timed region, one device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (frequencies and shot panels would be "
              "split across GPUs with one gradient all-reduce; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch

    import germaine_amd as G
    from germaine_amd import fdfd, models

    L = G.lib()
    w = models.config(args.config)
    nf, ns = len(w.freqs), len(w.src)
    log(f"bench: {w.name} {w.nx}x{w.ny} freqs={w.freqs} shots={ns}")
    m1 = torch.as_tensor(w.m1, device="cuda")
    m2 = torch.as_tensor(w.m2, device="cuda")
    ksrc = torch.as_tensor(np.asarray(w.src, np.int32), device="cuda")
    cfg = dict(small_max=args.small_max, rw_cap=args.fused_rw, fused_mmax=args.fused_mmax,
               refine_layout=args.refine_layout, persistent=args.persistent, deep_min_m=args.deep_min_m)
    fdfd.solve_tuning(cfg["rw_cap"], cfg["fused_mmax"], cfg["refine_layout"])
    pl = fdfd.Plan(w.nx, w.ny, small_max=cfg["small_max"])
    n = pl.n
    st = torch.cuda.current_stream()
    amps = [fdfd.source_amp(w.physics, f, w.s, w.dh) for f in w.freqs]
    Bs = torch.empty((nf, ns, n), dtype=torch.complex128, device="cuda")
    Xs = torch.empty((nf, ns, n), dtype=torch.complex128, device="cuda")
    Xs2 = torch.empty((nf, ns, n), dtype=torch.complex128, device="cuda")
    freqs = list(range(nf))

    store = {}
    vbuf = torch.empty((nf, pl.nnz), dtype=torch.complex128, device="cuda")

    def step(precision, refine):
        vals = pl.assemble(w.physics, w.dh, w.freqs, m1, m2, s=w.s, npml=w.npml, a0_pml=w.a0_pml, out=vbuf)
        # factor storage is allocated once per precision (the warm-up step) and refactored in place
        # afterwards, as one replay of the captured factorisation graph
        if precision not in store:
            F = pl.factor(vals, precision)
        elif args.no_graph:
            F = store[precision].refactor(vals)
        else:
            F = store[precision].refactor_graph(vals)
        store[precision] = F
        for f in range(nf):
            G.check(L.gfdfd_spike_panel(n, ns, ctypes.c_void_p(ksrc.data_ptr()), amps[f].real, amps[f].imag,
                                        ctypes.c_void_p(Bs[f].data_ptr()), ctypes.c_void_p(st.cuda_stream)))
        if refine:
            _, it, rr = F.solve_refine_batch(freqs, Bs, tol=args.tol, max_iter=20, out=Xs)
            its = list(zip(it, rr))
        else:
            F.solve_batch(freqs, Bs, out=Xs)
            its = []
        return F, its

    def profiling(on, cls=None):
        if on:
            G.check(L.gfdfd_profile_select(cls.encode() if cls else None))
            L.gfdfd_profile_reset()
        L.gfdfd_profile(1 if on else 0)

    def timed(precision, refine, steps, warmup, profile=False, legs=False):
        """Warm-up steps -- the first one, whose factorisation runs eagerly (the
        later ones replay the captured graph, which records no events), with
        every kernel class event-timed when profile: the per-kernel table of one
        step and the dominant class -- then the timed steps, in which only the
        dominant class records events (two per launch), so the headline time
        carries no per-launch profiling of the other classes.  With legs, after
        the timed steps: the factorisation alone (graph replays) and one
        batched solve pass alone (no refinement), each the median of 5."""
        kw, dom = {}, None
        for i in range(warmup):
            if profile and i == 0:
                profiling(True)
            step(precision, refine)
            if profile and i == 0:
                torch.cuda.synchronize()
                profiling(False)
                kw = kernel_stats(L, ctypes)
                dom = max(kw, key=lambda k: kw[k]["ms"]) if kw else None
        if warmup and not args.no_graph:
            store[precision].refactor_graph(vbuf)   # capture the factorisation graph before the timed steps
        torch.cuda.synchronize()
        if profile:
            profiling(True, dom)
        t0 = time.perf_counter()
        for _ in range(steps):
            _, its = step(precision, refine)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        kt = {}
        if profile:
            profiling(False)
            kt = kernel_stats(L, ctypes)
            G.check(L.gfdfd_profile_select(None))
        leg = {}
        if legs and precision in store:
            F = store[precision]
            tf, tsv = [], []
            for _ in range(5):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                F.refactor_graph(vbuf) if not args.no_graph else F.refactor(vbuf)
                torch.cuda.synchronize()
                tf.append(time.perf_counter() - t1)
            for _ in range(5):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                F.solve_batch(freqs, Bs, out=Xs2)
                torch.cuda.synchronize()
                tsv.append(time.perf_counter() - t1)
            leg = dict(factor_s=float(np.median(tf)), solve_pass_s=float(np.median(tsv)))
        store.pop(precision, None)                  # release this precision's factors
        return t, its, kw, (kt if dom else {}), leg

    # --- kernel selection: autotuner (before any timed region) ----------------------------
    chosen_by = "flags"
    if args.tune != "off":
        from germaine_amd import autotune

        def measure(c):
            nonlocal pl
            fdfd.solve_tuning(c["rw_cap"], c["fused_mmax"], c["refine_layout"])
            pl = fdfd.Plan(w.nx, w.ny, small_max=c["small_max"])
            return timed(G.FP32, True, 1, 1)[0]

        k = autotune.key(w.nx, w.ny, nf, ns, G.FP32, torch.cuda.get_device_name(0))
        cfg, chosen_by, _ = autotune.tune(k, measure, log=log, use_cache=args.tune == "auto")
        log(f"autotune ({chosen_by}): {cfg}")
        fdfd.solve_tuning(cfg["rw_cap"], cfg["fused_mmax"], cfg["refine_layout"])
        pl = fdfd.Plan(w.nx, w.ny, small_max=cfg["small_max"])

        # second stage: the factorisation schedule, timed on eager refactorisations
        def measure_factor(c, precision=G.FP32):
            fdfd.factor_tuning(0, c["persistent"], c["deep_min_m"])
            vals = pl.assemble(w.physics, w.dh, w.freqs, m1, m2, s=w.s, npml=w.npml, a0_pml=w.a0_pml, out=vbuf)
            F = pl.factor(vals, precision)
            ts = []
            for _ in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                F.refactor(vals)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            del F
            return float(np.median(ts))

        fcfg, fchosen, _ = autotune.tune(k + "|factor", measure_factor, cands=autotune.factor_candidates(), log=log,
                                         use_cache=args.tune == "auto")
        log(f"autotune factor schedule ({fchosen}): {fcfg}")
        cfg.update(fcfg)
        chosen_by = chosen_by if chosen_by == fchosen else f"{chosen_by}+{fchosen}"
    fdfd.factor_tuning(0, cfg["persistent"], cfg["deep_min_m"])
    info = pl.info

    # --- headline: FP32 fronts + FP64 refinement, batched -----------------------------
    t32, its32, kst, kdom, leg32 = timed(G.FP32, True, args.steps, max(1, args.warmup), profile=True, legs=True)
    log(f"fp32+IR batched: {t32 * 1e3:.1f} ms/step, refinement (iters, relres) per freq {its32}, legs {leg32}")
    X32 = Xs.clone() if not args.no_check else None

    # --- FP64 fronts, batched, no refinement ---------------------------------------------
    n64 = max(1, args.steps // 2)
    t64, kst64 = None, {}
    leg64 = {}
    cfg64 = dict(persistent=cfg["persistent"], deep_min_m=cfg["deep_min_m"], chosen_by="as fp32+IR")
    if not args.no_fp64:
        if args.tune != "off":
            # the FP64 factorisation's own schedule (its MFMA rate and LDS budget differ)
            k64 = autotune.key(w.nx, w.ny, nf, ns, G.FP64, torch.cuda.get_device_name(0)) + "|factor"
            c64, how, _ = autotune.tune(k64, lambda c: measure_factor(c, G.FP64), cands=autotune.factor_candidates(),
                                        log=log, use_cache=args.tune == "auto")
            cfg64 = dict(c64, chosen_by=how)
            log(f"autotune fp64 factor schedule ({how}): {c64}")
        fdfd.factor_tuning(0, cfg64["persistent"], cfg64["deep_min_m"])
        t64, _, kst64, _, leg64 = timed(G.FP64, False, n64, 1, profile=True, legs=True)
        fdfd.factor_tuning(0, cfg["persistent"], cfg["deep_min_m"])
        log(f"fp64 batched: {t64 * 1e3:.1f} ms/step, legs {leg64}")
    err = None
    if X32 is not None and t64 is not None:
        err = []
        for f in range(nf):
            d = torch.linalg.vector_norm(X32[f] - Xs[f], dim=1) / torch.linalg.vector_norm(Xs[f], dim=1)
            err.append(float(d.max()))
        log(f"fp32+IR vs fp64 rel wavefield error per freq: {err}")
    X32_last = X32[nf - 1].clone() if X32 is not None else None
    del X32

    # --- plain FP64: what the drop-in executes (per-frequency factor, one solve per shot) ---
    tplain, t1solve = None, None
    if not args.no_plain:
        fdfd.factor_tuning(0, cfg64["persistent"], cfg64["deep_min_m"])   # FP64 factors: the FP64 schedule
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tsolves = 0.0
        for f in range(nf):
            vals = pl.assemble(w.physics, w.dh, [w.freqs[f]], m1, m2, s=w.s, npml=w.npml, a0_pml=w.a0_pml)
            F = pl.factor(vals, G.FP64)
            G.check(L.gfdfd_spike_panel(n, ns, ctypes.c_void_p(ksrc.data_ptr()), amps[f].real, amps[f].imag,
                                        ctypes.c_void_p(Bs[f].data_ptr()), ctypes.c_void_p(st.cuda_stream)))
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for s in range(ns):
                F.solve(0, Bs[f][s], out=Xs[f][s])
            torch.cuda.synchronize()
            tsolves += time.perf_counter() - t1
            del F
        torch.cuda.synchronize()
        tplain = time.perf_counter() - t0
        t1solve = tsolves / (nf * ns)
        log(f"plain fp64 per-shot: {tplain * 1e3:.1f} ms, {t1solve * 1e3:.3f} ms per single-RHS solve")

    if args.copy_probe:
        # a plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # calibrate the HBM-traffic counters of the solve kernels
        src = torch.empty(args.copy_probe << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # --- roofline of the dominant kernel, event-timed over the timed steps ------------------
    # (kst: every class over the first warm-up step; kdom: the dominant class over the timed steps)
    src = kdom if kdom else kst
    dom = max(src, key=lambda k: src[k]["ms"]) if src else None
    roof = None
    if dom:
        roof = roofline(dom, src[dom], PEAK["fp32_mfma"])
        roof["launches_per_step"] = src[dom]["launches"] / args.steps
        roof["ms_per_step"] = src[dom]["ms"] / args.steps
        # HBM bytes per launch of this kernel from the committed FETCH_SIZE/WRITE_SIZE
        # pmc runs of this bench (copy-calibrated); null when they cover another kernel
        tr = measured_traffic(dom, w.name)
        if tr is not None:
            roof["traffic"] = tr["traffic_bytes_per_launch"]
            roof["traffic_source"] = tr["source"]

    # --- CPU baseline ---------------------------------------------------------------------------
    cpu, err_oracle = None, None
    if not args.no_cpu:
        try:
            from oracle import cpu_baseline
            cpu, idx, Xo = cpu_baseline.measure(w, threads=args.cpu_threads or None)
            if X32_last is not None:
                # the headline path's wavefields against the oracle's FP64 solutions of the same shots
                Xd = X32_last[idx].cpu().numpy()
                err_oracle = max(float(np.linalg.norm(Xd[r] - Xo[r]) / np.linalg.norm(Xo[r]))
                                 for r in range(len(idx)))
                log(f"fp32+IR vs oracle ({len(idx)} shots at {w.freqs[-1]:g} Hz): {err_oracle:.2e}")
        except Exception as e:                       # reported, never fatal
            log(f"cpu baseline unavailable: {e}")

    value = nf * ns / t32
    es32, es64 = 8.0, 16.0
    fac_flops = nf * info.factor_flops
    sweep_flops = 8.0 * info.nnz_lu * ns * nf
    def legs_of(leg, es, peak):
        if not leg:
            return None
        by = (info.nnz_lu * es + 2.0 * n * ns * es) * nf
        return {"factor_ms": round(leg["factor_s"] * 1e3, 3),
                "factor_tflops": round(fac_flops / leg["factor_s"] / 1e12, 2),
                "factor_mfma_frac": round(fac_flops / leg["factor_s"] / peak, 4),
                "solve_pass_ms": round(leg["solve_pass_s"] * 1e3, 3),
                "solve_tflops": round(sweep_flops / leg["solve_pass_s"] / 1e12, 2),
                "solve_mfma_frac": round(sweep_flops / leg["solve_pass_s"] / peak, 4),
                "solve_gbps": round(by / leg["solve_pass_s"] / 1e9, 1),
                "solve_hbm_frac": round(by / leg["solve_pass_s"] / PEAK["hbm"], 4)}
    res = {
        "metric": "shot-solves/sec (factor+solve) FP32+IR vs FP64 LU at rel err <=1e-8; MFMA %",
        "value": round(value, 2),
        "unit": "shot-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 fronts + f64 refinement",
        "data": "synthetic seeded layered-fault model, point sources",
        "config": {"workload": f"{w.name}: AC {w.nx}x{w.ny} incl. {w.npml}-cell PML, h={w.dh} m, "
                               f"{nf} freqs {w.freqs[0]:g}-{w.freqs[-1]:g} Hz batched, {ns}-shot panel",
                   "n": n, "nnz": pl.nnz, "fronts": info.fronts, "factor_gflop_per_freq": info.factor_flops / 1e9,
                   "nnz_lu": info.nnz_lu, "refine_tol": args.tol,
                   "tuning": dict(cfg, chosen_by=chosen_by)},
        "legs": legs_of(leg32, es32, PEAK["fp32_mfma"]),
        "legs_note": "factor_*: one batched factorisation of all frequencies (graph replay), against the dense "
                     "MFMA peak of the factor precision; solve_*: one batched forward+backward solve pass of all "
                     "panels without refinement (incl. the caller-layout transposes), flops 8 nnz(L+U) per RHS, "
                     "bytes nnz(L+U) + 2 n nrhs entries",
        "fp64_batched": None if t64 is None else {
            "ms_per_step": round(t64 * 1e3, 3), "value": round(nf * ns / t64, 2),
            "legs": legs_of(leg64, es64, PEAK["fp64_mfma"]),
            "tuning": cfg64,
            "kernels_ms_per_step": {k: round(v["ms"], 3) for k, v in kst64.items()}},
        "plain_fp64_per_shot": None if tplain is None else {
            "ms": round(tplain * 1e3, 3), "value": round(nf * ns / tplain, 2),
            "ms_per_single_rhs_solve": round(t1solve * 1e3, 4),
            "solve_hbm_frac": round(info.nnz_lu * es64 / t1solve / PEAK["hbm"], 4)},
        "speedup_vs_plain_fp64": None if tplain is None else round(tplain / t32, 3),
        "refinement": [{"iters": a, "relres": b} for a, b in its32],
        "max_rel_err_vs_fp64": None if err is None else max(err),
        "max_rel_err_vs_oracle": err_oracle,
        "roofline": roof,
        "kernels_note": "every kernel class event-timed over the first warm-up step (eager factorisation)",
        "kernels": {k: {"ms_per_step": round(v["ms"], 3), "launches_per_step": v["launches"],
                        "tflops": round(v["flops"] / max(v["ms"], 1e-9) * 1e-9, 3),
                        "gbps": round(v["bytes"] / max(v["ms"], 1e-9) * 1e-6, 1)} for k, v in kst.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
