"""Autotuner of the solve path's kernel selection (no counterpart in the reference).

Switches it chooses between (all correct, their speed depends on the workload):

* ``small_max``: largest front order factored wholly in LDS by one workgroup (plan option; the
  rest take the blocked panel / TRSM / MFMA-update path);
* ``rw_cap``: widest right-hand-side tile of the fused per-front solve kernels (16, 32, 64);
* ``fused_mmax``: largest front order of a multi-block level that takes the fused per-front
  solve kernels instead of the per-block ones (0: none, -1: as far as LDS allows);
* ``refine_layout``: the refinement in row-major grid order with one fused residual-and-update
  pass per step (0), or in the caller's layout with a transpose per correction (1);

and, in a second stage with the solve switches fixed, the factorisation's schedule:

* ``persistent``: the small subtrees in one persistent launch (1) or level by level (0);
* ``deep_min_m``: the front-size bucket boundary of the blocked path -- levels whose largest
  blocked front has at least this order factor two 32-column panels per step and update the
  trailing block once per 64 pivots (product depth 64), smaller ones once per panel (depth 32).

``tune`` times a caller-supplied measurement (seconds per step of the real workload for one
configuration) over the candidates, keeps the fastest and records it in a JSON cache keyed by the
workload shape and the device, so later runs look the choice up instead of searching again.
"""
import json
import os

CACHE = os.environ.get("GFDFD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "gfdfd", "tuning.json"))

DEFAULT = dict(small_max=64, rw_cap=64, fused_mmax=0, refine_layout=0)
NO_DEEP = 1 << 30
FACTOR_DEFAULT = dict(persistent=1, deep_min_m=0)


def candidates():
    """The search space: every combination below (16 configurations)."""
    out = []
    for small_max in (64, 32):
        for rw_cap in (64, 32):
            for fused_mmax in (0, -1):
                for refine_layout in (0, 1):
                    out.append(dict(small_max=small_max, rw_cap=rw_cap, fused_mmax=fused_mmax,
                                    refine_layout=refine_layout))
    return out


def factor_candidates():
    """The factorisation schedules (8 configurations): product depth 64 on every blocked level,
    on levels with fronts of order >= 384 or >= 1024 only, or on none; subtrees persistent or not."""
    return [dict(persistent=p, deep_min_m=d) for p in (1, 0) for d in (0, 384, 1024, NO_DEEP)]


def key(nx, ny, nfreq, nrhs, precision, device):
    return f"{device}|{nx}x{ny}|f{nfreq}|r{nrhs}|p{precision}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(k, path=CACHE):
    rec = _load(path).get(k)
    return None if rec is None else dict(rec["best"])


def store(k, best, table, path=CACHE):
    data = _load(path)
    data[k] = dict(best=best, table=table)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, path)


def tune(k, measure, cands=None, log=None, path=CACHE, use_cache=True):
    """Return (best configuration, how it was chosen, table of (configuration, seconds)).

    measure(cfg) -> seconds per step of the workload run with cfg (the caller warms up and
    times; a configuration that raises is skipped and reported)."""
    if use_cache:
        best = lookup(k, path)
        if best is not None:
            return best, "cache", []
    table = []
    for cfg in (cands or candidates()):
        try:
            t = float(measure(cfg))
        except Exception as e:          # a failing candidate is reported, never chosen
            if log:
                log(f"autotune: {cfg} failed: {e}")
            continue
        table.append((cfg, t))
        if log:
            log(f"autotune: {cfg} -> {t * 1e3:.1f} ms")
    if not table:
        return dict((cands or candidates())[0]), "default", table
    best = min(table, key=lambda e: e[1])[0]
    store(k, best, [dict(cfg=c, seconds=t) for c, t in table], path)
    return dict(best), "search", table
