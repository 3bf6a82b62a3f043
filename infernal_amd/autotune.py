"""Launch-parameter autotuner for the standard scanners.

The scanners pick their launch shape from occupancy rules: waves per
workgroup, chunk length (at least 2W, else one chunk per resident
wavefront), where the local-begin maxima live (LDS or the wavefront's global
slot) and, for CYK, the four-block wide kernel.  Those rules are right for
the common shapes; for a given set of models and windows the fastest
choice can differ.  tune() times the candidates on a sample of the windows
and keeps the fastest, one parameter at a time (coordinate descent from the
automatic choice), and caches the result per workload key.  apply() sets a
configuration for the calls made inside it.

Every candidate computes the same cells and the same hits; only the time
differs, so tuning never changes results.
"""
import contextlib
import os
import time

import numpy as np

# the knobs, each with its automatic value first
KNOBS = {
    "wpb": [0, 16, 8, 4],                 # waves per workgroup (0: automatic)
    "chunk_w": [0, 2, 4, 8],              # chunk length in multiples of the largest W (0: automatic)
    "begin_acc": ["auto", "lds", "global"],
    "wide": ["auto", "0", "1"],           # CYK only
}
_CACHE = {}


@contextlib.contextmanager
def apply(config, w_max):
    """Run the enclosed scans with `config` (a dict over KNOBS); w_max is the
    largest W of the models scanned (for chunk_w)."""
    import infernal_amd as ia
    old = {k: os.environ.get(k) for k in ("INFX_BEGIN_ACC", "INFX_WIDE")}
    try:
        for key, env in (("begin_acc", "INFX_BEGIN_ACC"), ("wide", "INFX_WIDE")):
            v = config.get(key, "auto")
            if v == "auto":
                os.environ.pop(env, None)
            else:
                os.environ[env] = v
        ia.set_launch_params(int(config.get("chunk_w", 0)) * int(w_max), int(config.get("wpb", 0)))
        yield
    finally:
        ia.set_launch_params(0, 0)
        for env, v in old.items():
            if v is None:
                os.environ.pop(env, None)
            else:
                os.environ[env] = v


def workload_key(algo, cms, qdbidx, windows):
    lens = np.array([w[1] - w[0] + 1 for w in windows])
    return (algo, qdbidx, tuple((cm.c.name.decode(), cm.W) for cm in cms), len(windows),
            int(np.log2(max(1, lens.sum()))))


def tune(algo, cms, qdbidx, dsq, windows, cutoffs, sample=0.1, seed=0, run=None, reps=1, knobs=("wpb", "chunk_w")):
    """Coordinate-descent tuning of `knobs` (names in KNOBS) on a sample of
    `windows` ((i, j, model) triples, as ScanBatch takes).  `run(config,
    windows)` times one call and returns seconds; by default it runs
    ia.ScanBatch under apply(config) (the batched launch has one kernel
    variant, so only the waves per workgroup and the chunk length apply).
    Returns (best config, [(config, seconds), ...])."""
    import infernal_amd as ia
    key = workload_key(algo, cms, qdbidx, windows) + (tuple(knobs),)
    if key in _CACHE:
        return _CACHE[key]
    w_max = max(cm.W for cm in cms)
    rng = np.random.default_rng(seed)
    k = max(1, int(round(sample * len(windows))))
    pick = np.sort(rng.choice(len(windows), k, replace=False))
    sub = [windows[i] for i in pick]

    def default_run(config, ws):
        with apply(config, w_max):
            t0 = time.perf_counter()
            ia.ScanBatch(algo, cms, qdbidx, dsq, ws, cutoffs)
            return time.perf_counter() - t0

    run = run or default_run
    trials = []

    def timed(cfg):
        try:
            t = min(run(cfg, sub) for _ in range(reps))
        except ia.ScanError:
            t = float("inf")                 # the shape does not fit this workload (LDS, registers)
        trials.append((dict(cfg), t))
        return t

    best = {kn: vals[0] for kn, vals in KNOBS.items()}
    if algo != "cyk":
        best["wide"] = "auto"
    run(best, sub[:1])                       # first-use set-up, not timed
    best_t = timed(best)
    for kn, vals in KNOBS.items():
        if kn not in knobs or (kn == "wide" and algo != "cyk"):
            continue
        for v in vals[1:]:
            cfg = dict(best)
            cfg[kn] = v
            t = timed(cfg)
            if t < best_t:
                best, best_t = cfg, t
    _CACHE[key] = (best, trials)
    return best, trials
