"""The launch autotuner's search and caching, with a stand-in timer (no GPU):
coordinate descent must find the fastest value of each knob from the
automatic start, skip the CYK-only knob for Inside, cache per workload, and
apply() must set and restore the launch overrides."""
import os

import infernal_amd as ia
from infernal_amd import autotune

CMDIR = os.path.join(os.path.dirname(__file__), "golden", "cm")


def _cms():
    return [ia.CM.read(os.path.join(CMDIR, n)).configure(local=True) for n in ("tRNA.c.cm", "Vault.c.cm")]


def test_coordinate_descent_finds_the_cheapest_knobs():
    cost = {"wpb": {0: 5, 16: 3, 8: 1, 4: 2}, "chunk_w": {0: 4, 2: 1, 4: 3, 8: 5},
            "begin_acc": {"auto": 2, "lds": 1, "global": 3}, "wide": {"auto": 2, "0": 3, "1": 1}}
    calls = []

    def run(cfg, ws):
        calls.append(len(ws))
        return sum(cost[k][cfg[k]] for k in cost)

    cms = _cms()
    wins = [(1 + 300 * k, 250 + 300 * k, k % 2) for k in range(40)]
    every = tuple(autotune.KNOBS)
    best, trials = autotune.tune("cyk", cms, 1, None, wins, [0.0, 0.0], sample=0.25, run=run, knobs=every)
    assert best == {"wpb": 8, "chunk_w": 2, "begin_acc": "lds", "wide": "1"}
    assert len(trials) == 1 + 3 + 3 + 2 + 2
    assert calls[0] == 1 and all(c == 10 for c in calls[1:])
    # cached per workload: no new calls
    n = len(calls)
    assert autotune.tune("cyk", cms, 1, None, wins, [0.0, 0.0], sample=0.25, run=run, knobs=every)[0] == best
    assert len(calls) == n
    # Inside: the CYK-only knob is left alone
    best_i, trials_i = autotune.tune("finside", cms, 2, None, wins, [0.0, 0.0], sample=0.25, run=run, knobs=every)
    assert best_i["wide"] == "auto" and all(t[0]["wide"] == "auto" for t in trials_i)
    # the batched launch's default knobs: waves per workgroup and chunk length only
    best_b, trials_b = autotune.tune("cyk", cms, 1, None, wins[:20], [0.0, 0.0], sample=0.5, run=run)
    assert len(trials_b) == 1 + 3 + 3 and best_b["wpb"] == 8 and best_b["chunk_w"] == 2
    assert best_b["begin_acc"] == "auto" and best_b["wide"] == "auto"


def test_apply_sets_and_restores_overrides():
    os.environ.pop("INFX_BEGIN_ACC", None)
    os.environ["INFX_WIDE"] = "0"
    with autotune.apply({"wpb": 4, "chunk_w": 2, "begin_acc": "global", "wide": "auto"}, w_max=100):
        assert os.environ.get("INFX_BEGIN_ACC") == "global"
        assert "INFX_WIDE" not in os.environ
    assert "INFX_BEGIN_ACC" not in os.environ and os.environ.get("INFX_WIDE") == "0"
    os.environ.pop("INFX_WIDE", None)


def test_candidates_that_do_not_fit_are_skipped():
    """A launch shape the scanner refuses (ScanError: LDS or registers do not
    fit the workload) is recorded as infinitely slow, and the search goes on."""
    def run(cfg, ws):
        if cfg["wpb"] == 16:
            raise ia.ScanError(7, "no workgroup shape fits")
        return {0: 3.0, 8: 1.0, 4: 2.0}[cfg["wpb"]] + {0: 2.0, 2: 1.0, 4: 3.0, 8: 4.0}[cfg["chunk_w"]]

    cms = _cms()
    wins = [(1 + 300 * k, 250 + 300 * k, k % 2) for k in range(30)]
    best, trials = autotune.tune("cyk", cms, 1, None, wins, [0.0, 0.0], sample=0.5, run=run)
    assert best["wpb"] == 8 and best["chunk_w"] == 2
    assert [t for c, t in trials if c["wpb"] == 16] == [float("inf")]
