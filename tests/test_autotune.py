"""Host logic of the autotuner: the fastest measured configuration wins, failing
candidates are skipped, the choice is cached per workload key and found again."""
from germaine_amd import autotune


def test_search_picks_fastest_and_caches(tmp_path):
    path = str(tmp_path / "tuning.json")
    cands = autotune.candidates()
    assert len(cands) == 16 and autotune.DEFAULT in cands

    def measure(c):
        if c["rw_cap"] == 32 and c["fused_mmax"] == -1:
            raise RuntimeError("not supported here")
        return 1.0 + 0.1 * c["refine_layout"] + (0.05 if c["small_max"] == 64 else 0.0) + 0.01 * (c["rw_cap"] == 64)

    k = autotune.key(101, 51, 3, 40, 32, "dev")
    best, how, table = autotune.tune(k, measure, cands, path=path)
    assert how == "search"
    assert best == dict(small_max=32, rw_cap=32, fused_mmax=0, refine_layout=0)
    assert len(table) == 12                       # four candidates raised
    assert autotune.lookup(k, path) == best
    again, how2, _ = autotune.tune(k, lambda c: 1 / 0, cands, path=path)
    assert how2 == "cache" and again == best
    other, how3, _ = autotune.tune(autotune.key(101, 51, 3, 41, 32, "dev"), measure, cands, path=path,
                                   use_cache=True)
    assert how3 == "search" and other == best


def test_all_candidates_failing_falls_back_to_default(tmp_path):
    path = str(tmp_path / "tuning.json")

    def measure(c):
        raise RuntimeError("no device")

    best, how, table = autotune.tune("k", measure, path=path)
    assert how == "default" and best == autotune.DEFAULT and table == []
    assert autotune.lookup("k", path) is None
