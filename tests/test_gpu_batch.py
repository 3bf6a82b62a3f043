"""infx_FastScanBatch: several models' windows of one target in one launch.

Each model's hits and best score must equal, bit for bit, those of
infx_FastScanWindows over that model's windows alone (itself checked against
the oracle in test_gpu_parity_std.py).  The models differ in size (tRNA,
the 350-nt multiloop, Cobalamin with bifurcations); windows of different
models overlap; a forced short chunk length gives many chunks per window.
"""
import os

import numpy as np
import pytest

import infernal_amd as ia
from helpers import CMDIR, load_cm, hits_equal, describe_hit_diff

pytestmark = pytest.mark.gpu


def _target(L, seed):
    from infernal_amd import synth
    frags = []
    for k, name in enumerate(("tRNA.c.cm", "synth_multiloop350.cm")):
        frags += synth.homologs(ia.CM.read(os.path.join(CMDIR, name)), 3, seed=seed + k)
    bg = synth.markov2_background(L, gc=0.45, seed=seed + 7)
    codes, _ = synth.plant(bg, frags, seed=seed + 9)
    return ia.dsq_from_codes(codes)


WINDOWS = {0: [(1, 5000), (6000, 9000), (15000, 15100)],
           1: [(2000, 12000), (12500, 14000)],
           2: [(500, 8000), (10000, 20000)]}


@pytest.mark.parametrize("algo", ["cyk", "finside", "iinside"])
def test_batch_equals_per_model_windows(algo):
    names = ["tRNA.c.cm", "synth_multiloop350.cm", "Cobalamin.c.cm"]
    cms = [load_cm(n) for n in names]
    dsq = _target(20000, seed=61)
    qdb = 1 if algo == "cyk" else 2
    cutoffs = [-5.0, 0.0, -2.0]
    wins = [(i, j, m) for m, ws in WINDOWS.items() for (i, j) in ws]
    print("case batch", algo, len(wins), flush=True)
    ia.set_launch_params(chunk_len=900)
    try:
        got = ia.ScanBatch(algo, cms, qdb, dsq, wins, cutoffs)
    finally:
        ia.set_launch_params(0, 0)
    key = lambda h: (h[0], h[1])
    for m, cm in enumerate(cms):
        want = ia.ScanWindows(algo, cm, qdb, dsq, WINDOWS[m], cutoffs[m])
        g = sorted(got[m]["hits"], key=key)
        w = sorted(want["hits"], key=key)
        assert hits_equal(g, w), (names[m], describe_hit_diff(g, w))
        assert np.float32(got[m]["sc"]).view(np.uint32) == np.float32(want["sc"]).view(np.uint32), (names[m], got[m]["sc"], want["sc"])
    assert sum(len(x["hits"]) for x in got) > 0


CFG4_MODELS = ["synth_cloverleaf76.cm", "tRNA.c.cm", "synth_multiloop350.cm", "Cobalamin.c.cm", "synth_nested1500.cm"]


def _cfg4_windows(cms, n_per, L, seed):
    """n_per windows of 1.2 W per model at seeded, per-model disjoint
    positions of a target of L residues (models' windows overlap freely)."""
    rng = np.random.default_rng(seed)
    wins = []
    for m, cm in enumerate(cms):
        wl = int(1.2 * cm.W)
        starts = np.sort(rng.choice(np.arange(1, L - wl, wl + 1), n_per, replace=False))
        wins += [(int(a), int(a) + wl - 1, m) for a in starts]
    return wins


@pytest.mark.parametrize("algo", ["cyk", "iinside"])
def test_batch_against_oracle_model_sizes(algo):
    """The many-models workload's shape against the oracle: five models of
    76 to 1,500 consensus nt (W 89 to 1,522) in one launch, windows of 1.2 W,
    a target of random background with each model's emitted homologs planted;
    every model's hits and best score equal the oracle's scans of its windows
    one by one (hits of the windows pooled, as a batched call returns them)."""
    import concurrent.futures as cf
    import oracle_bind as ob
    from infernal_amd import synth
    cms = [load_cm(n) for n in CFG4_MODELS]
    frags = []
    for k, n in enumerate(CFG4_MODELS):
        frags += synth.homologs(ia.CM.read(os.path.join(CMDIR, n)), 2, seed=700 + k)
    L = 60000
    bg = synth.markov2_background(L, gc=0.45, seed=707)
    codes, _ = synth.plant(bg, frags, seed=709)
    dsq = ia.dsq_from_codes(codes)
    n_per = 3 if algo == "cyk" else 2
    wins = _cfg4_windows(cms, n_per, L, seed=711)
    if algo == "iinside":   # the 1,500-nt model's Inside takes the oracle ~15 s per window: one
        wins = [w for w in wins if w[2] != 4 or w[0] == min(x[0] for x in wins if x[2] == 4)]
    qdb = 1 if algo == "cyk" else 2
    cutoffs = [0.0] * len(cms)
    print("case batch-vs-oracle", algo, [(w[2], w[1] - w[0] + 1) for w in wins], flush=True)
    got = ia.ScanBatch(algo, cms, qdb, dsq, wins, cutoffs)
    oalgo = ob.ORACLE_CYK if algo == "cyk" else ob.ORACLE_IINSIDE
    for cm in cms:
        if algo == "cyk":
            cm.c.search_opts &= ~ia.CM_SEARCH_INSIDE
        else:
            cm.c.search_opts |= ia.CM_SEARCH_INSIDE
    with cf.ThreadPoolExecutor(8) as ex:
        outs = list(ex.map(lambda w: ob.scan(cms[w[2]], oalgo, qdb, dsq, w[0], w[1], 0.0), wins))
    key = lambda h: (h[0], h[1])
    for m, name in enumerate(CFG4_MODELS):
        mine = [o for w, o in zip(wins, outs) if w[2] == m]
        want = sorted([h for o in mine for h in o["hits"]], key=key)
        g = sorted(got[m]["hits"], key=key)
        assert hits_equal(g, want), (name, describe_hit_diff(g, want))
        best = max(o["sc"] for o in mine)
        assert np.float32(got[m]["sc"]).view(np.uint32) == np.float32(best).view(np.uint32), (name, got[m]["sc"], best)
    assert sum(len(x["hits"]) for x in got) > 0


def test_batch_graph_replays_equal_direct_launches():
    """The batched launch through a captured graph: three replays of one
    launch shape, with direct launches and host synchronisations between
    them, then a different shape (a new capture) and a replay of it; every
    result bit-identical to the direct launch of the same call."""
    names = ["tRNA.c.cm", "synth_multiloop350.cm", "Cobalamin.c.cm"]
    cms = [load_cm(n) for n in names]
    dsq = _target(20000, seed=63)
    cutoffs = [-5.0, 0.0, -2.0]
    wins_a = [(i, j, m) for m, ws in WINDOWS.items() for (i, j) in ws]
    wins_b = [w for w in wins_a if w[2] != 1]
    key = lambda h: (h[0], h[1])

    def same(x, y):
        for m in range(len(cms)):
            assert hits_equal(sorted(x[m]["hits"], key=key), sorted(y[m]["hits"], key=key)), names[m]
            assert np.float32(x[m]["sc"]).view(np.uint32) == np.float32(y[m]["sc"]).view(np.uint32), names[m]

    direct_a = ia.ScanBatch("cyk", cms, 1, dsq, wins_a, cutoffs)
    direct_b = ia.ScanBatch("cyk", cms, 1, dsq, wins_b, cutoffs)
    try:
        for rep, wins, want in ((0, wins_a, direct_a), (1, wins_a, direct_a), (2, wins_a, direct_a),
                                (3, wins_b, direct_b), (4, wins_b, direct_b)):
            print("case batch graph replay", rep, len(wins), flush=True)
            ia.set_batch_graph(True)
            got = ia.ScanBatch("cyk", cms, 1, dsq, wins, cutoffs)
            ia.set_batch_graph(False)
            same(got, want)
            same(ia.ScanBatch("cyk", cms, 1, dsq, wins, cutoffs), want)   # a direct launch between replays
    finally:
        ia.set_batch_graph(False)
