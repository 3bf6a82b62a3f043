"""GPU standard scanners vs the oracle, through the C ABI.

Bar: FastCYKScan and FastIInsideScan bit-exact (scores, hit lists, envelopes,
per-state maxima); FastFInsideScan within 0.01 bits (it is bit-exact in
practice, because the GPU follows the same FLogsum order and table).
"""
import os

import numpy as np
import pytest

import infernal_amd as ia
import oracle_bind as ob
from helpers import CMDIR, load_cm, random_dsq, hits_equal, describe_hit_diff

pytestmark = pytest.mark.gpu

FINSIDE_TOL = 0.01


def _run(fn_gpu, algo, cm, qdb, dsq, i0, j0, cutoff, null3=False, greedy=True, env=0.0, vsc=False, tol=0.0):
    print("case", fn_gpu.__name__, cm.c.name.decode(), qdb, i0, j0, cutoff, null3, greedy, flush=True)
    g = fn_gpu(cm, qdb, dsq, i0, j0, cutoff, do_null3=null3, env_cutoff=env, want_vsc=vsc, greedy=greedy)
    o = ob.scan(cm, algo, qdb, dsq, i0, j0, cutoff, do_null3=null3, greedy=greedy, env_cutoff=env)
    if tol == 0.0:
        assert np.float32(g["sc"]).view(np.uint32) == np.float32(o["sc"]).view(np.uint32), (g["sc"], o["sc"])
        assert hits_equal(g["hits"], o["hits"]), describe_hit_diff(g["hits"], o["hits"])
        if vsc:
            assert np.array_equal(g["vsc"].view(np.uint32), o["vsc"].view(np.uint32))
    else:
        assert abs(g["sc"] - o["sc"]) <= tol
        assert len(g["hits"]) == len(o["hits"])
        for x, y in zip(g["hits"], o["hits"]):
            assert x[:4] == y[:4] and abs(x[4] - y[4]) <= tol
    assert (g["envi"], g["envj"]) == (o["envi"], o["envj"])
    return g, o


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Vault.c.cm", "5S_rRNA.c.cm"])
@pytest.mark.parametrize("qdb", [1, 2])
def test_cyk_local(name, qdb):
    cm = load_cm(name)
    dsq = random_dsq(3000, seed=11 + qdb)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, qdb, dsq, 1, 3000, 0.0, env=5.0, vsc=True)


def test_cyk_glocal_and_noqdb():
    cm = load_cm("tRNA.c.cm", local=False)
    dsq = random_dsq(1200, seed=5)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 0, dsq, 1, 1200, -20.0, vsc=True)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 1200, -20.0)


def test_cyk_chunked_long():
    """Many chunks: the W-1 warm-up must reproduce the reference cells."""
    cm = load_cm("tRNA.c.cm")
    dsq = random_dsq(20000, seed=3)
    ia.set_launch_params(chunk_len=500)
    try:
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 20000, 0.0, env=3.0)
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 777, 19000, 0.0, env=3.0)
    finally:
        ia.set_launch_params(0, 0)


def test_cyk_wide_kernel_chunked():
    """The wide-band CYK kernel over many short chunks, a window start that
    is not a chunk boundary, and short targets."""
    os.environ["INFX_WIDE"] = "1"
    ia.set_launch_params(chunk_len=500)
    try:
        cm = load_cm("tRNA.c.cm")
        dsq = random_dsq(20000, seed=4)
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 20000, 0.0, env=3.0, vsc=True)
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 777, 19000, 0.0, env=3.0)
        ia.set_launch_params(0, 0)
        for L in (1, 7, 218, 219, 1000):
            d2 = random_dsq(L, seed=200 + L)
            _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, d2, 1, L, -50.0, vsc=True)
        cm = load_cm("rnaseP-eubact.c.cm")
        d3 = random_dsq(3000, seed=61)
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, d3, 1, 3000, 0.0, vsc=True)
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 2, d3, 1, 3000, 0.0, null3=True, greedy=False)
    finally:
        os.environ.pop("INFX_WIDE", None)
        ia.set_launch_params(0, 0)


def test_cyk_null3_and_gamma():
    cm = load_cm("Vault.c.cm")
    dsq = random_dsq(4000, seed=9, gc=0.6)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 4000, 0.0, null3=True)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 4000, 0.0, greedy=False)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 4000, 0.0, null3=True, greedy=False)


def test_short_targets():
    cm = load_cm("tRNA.c.cm")
    for L in (1, 2, 7, 50, 217, 218, 219):
        dsq = random_dsq(L, seed=100 + L)
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, L, -50.0, vsc=True)
        _run(ia.FastIInsideScan, ob.ORACLE_IINSIDE, cm, 1, dsq, 1, L, -50.0, vsc=True)


def test_degenerate_residues():
    cm = load_cm("5S_rRNA.c.cm")
    dsq = random_dsq(2500, seed=21, degenerate=0.05)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 2500, 0.0)
    _run(ia.FastIInsideScan, ob.ORACLE_IINSIDE, cm, 2, dsq, 1, 2500, 0.0)


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Vault.c.cm"])
def test_finside(name):
    cm = load_cm(name)
    dsq = random_dsq(3000, seed=31)
    _run(ia.FastFInsideScan, ob.ORACLE_FINSIDE, cm, 2, dsq, 1, 3000, 0.0, env=5.0, vsc=True)
    _run(ia.FastFInsideScan, ob.ORACLE_FINSIDE, cm, 2, dsq, 1, 3000, 0.0, env=5.0, tol=FINSIDE_TOL)


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Vault.c.cm"])
def test_iinside(name):
    cm = load_cm(name)
    dsq = random_dsq(3000, seed=41)
    _run(ia.FastIInsideScan, ob.ORACLE_IINSIDE, cm, 2, dsq, 1, 3000, 0.0, env=5.0, vsc=True)
    _run(ia.FastIInsideScan, ob.ORACLE_IINSIDE, cm, 1, dsq, 1, 3000, 0.0, null3=True)


def test_bifurcating_models():
    for name in ("Cobalamin.c.cm", "rnaseP-eubact.c.cm"):
        cm = load_cm(name)
        dsq = random_dsq(2500, seed=51)
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 2500, 0.0, vsc=True)
        _run(ia.FastFInsideScan, ob.ORACLE_FINSIDE, cm, 2, dsq, 1, 1500, 0.0)


def _planted(cm_name, L, nhom, seed):
    """Markov background (40% GC) with emitted homologs of the named model."""
    from infernal_amd import synth
    frags = synth.homologs(ia.CM.read(os.path.join(CMDIR, cm_name)), nhom, seed=seed)
    bg = synth.markov2_background(L, gc=0.40, seed=seed + 1)
    codes, _ = synth.plant(bg, frags, seed=seed + 2)
    return ia.dsq_from_codes(codes)


@pytest.mark.parametrize("name", ["synth_cloverleaf76.cm", "synth_multiloop350.cm"])
def test_synthetic_models(name):
    """The benchmark's synthetic models on targets with planted homologs:
    CYK (QDB1) hit lists bit-exact, FInside (QDB2) within 0.01 bits."""
    cm = load_cm(name)
    dsq = _planted(name, 6000, 4, seed=21)
    g, _ = _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 6000, 0.0, env=5.0, vsc=True)
    assert g["sc"] > 20.0 and len(g["hits"]) >= 2
    _run(ia.FastFInsideScan, ob.ORACLE_FINSIDE, cm, 2, dsq, 1, 4000, 0.0, tol=FINSIDE_TOL)


@pytest.mark.parametrize("placement", ["auto", "lds"])
def test_synthetic_1500nt_cyk(placement, monkeypatch):
    """The 1,500-nt model (W 1522): CYK over a target spanning several chunks,
    with the local-begin accumulators where the launch puts them (the global
    slot, for this root band) and forced into LDS."""
    if placement != "auto":
        monkeypatch.setenv("INFX_BEGIN_ACC", placement)
    cm = load_cm("synth_nested1500.cm")
    dsq = _planted("synth_nested1500.cm", 5000, 2, seed=31)
    g, _ = _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 5000, 0.0, vsc=True)
    assert g["sc"] > 50.0


@pytest.mark.parametrize("placement,wide", [("global", "0"), ("lds", "0"), ("global", "1"), ("lds", "1")])
def test_begin_accumulator_placements(placement, wide, monkeypatch):
    """Local-begin maxima in the wavefront's global slot (int-keyed atomic
    max) or in LDS, and the wide-band CYK kernel (four blocks per step) or
    the narrow one, give the same bit-exact results for all three scanners."""
    monkeypatch.setenv("INFX_BEGIN_ACC", placement)
    monkeypatch.setenv("INFX_WIDE", wide)
    cm = load_cm("synth_multiloop350.cm")
    dsq = _planted("synth_multiloop350.cm", 6000, 4, seed=23)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 6000, 0.0, env=5.0, vsc=True)
    _run(ia.FastFInsideScan, ob.ORACLE_FINSIDE, cm, 2, dsq, 1, 4000, 0.0, tol=FINSIDE_TOL)
    _run(ia.FastIInsideScan, ob.ORACLE_IINSIDE, cm, 2, dsq, 1, 4000, 0.0, vsc=True)
    cm = load_cm("Cobalamin.c.cm")
    dsq = random_dsq(2500, seed=57)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 1, 2500, -5.0, vsc=True, greedy=False)
    _run(ia.FastIInsideScan, ob.ORACLE_IINSIDE, cm, 1, dsq, 1, 2500, 0.0, null3=True)


@pytest.mark.parametrize("algo", ["cyk", "finside"])
def test_scan_windows_equals_one_scan_per_window(algo):
    """infx_FastScanWindows: windows of 40..3000 residues (some shorter than
    W) in one launch return exactly the hits of one scan per window, and the
    oracle's."""
    cm = load_cm("synth_multiloop350.cm")
    dsq = _planted("synth_multiloop350.cm", 20000, 6, seed=41)
    wins = [(5, 44), (100, 900), (1200, 4200), (4300, 4310), (6000, 7999), (9000, 12000), (15000, 15500), (19500, 20000)]
    qdb, fn, oa = (1, ia.FastCYKScan, ob.ORACLE_CYK) if algo == "cyk" else (2, ia.FastFInsideScan, ob.ORACLE_FINSIDE)
    print("case windows", algo, len(wins), flush=True)
    ia.set_launch_params(chunk_len=500)    # several chunks per long window
    try:
        got = ia.ScanWindows(algo, cm, qdb, dsq, wins, -5.0)
    finally:
        ia.set_launch_params(0, 0)
    want, best = [], -1e30
    for (i, j) in wins:
        r = fn(cm, qdb, dsq, i, j, -5.0)
        o = ob.scan(cm, oa, qdb, dsq, i, j, -5.0)
        assert hits_equal(r["hits"], o["hits"]), describe_hit_diff(r["hits"], o["hits"])
        want += r["hits"]
        best = max(best, r["sc"])
    key = lambda h: (h[0], h[1])
    assert hits_equal(sorted(got["hits"], key=key), sorted(want, key=key))
    assert np.float32(got["sc"]).view(np.uint32) == np.float32(best).view(np.uint32)


def test_pinned_target():
    """A pinned target (infx_target_Pin): block scans read the device copy in
    place, with real residues beyond each block's ends instead of zero pads,
    and give the oracle's results for the block; the target's last block
    reads into the pin's trailing pad.  Windowed scans of the pinned buffer
    agree too.  Unpinned, the same calls upload their ranges again."""
    cm = load_cm("tRNA.c.cm")
    L = 9000
    dsq = random_dsq(L, seed=41)
    ia.pin_target(dsq)
    try:
        for (a, b) in ((1, 2500), (2301, 6000), (5801, L)):
            _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, a, b, 0.0, env=3.0, vsc=True)
            _run(ia.FastFInsideScan, ob.ORACLE_FINSIDE, cm, 2, dsq, a, b, 0.0, env=3.0)
        wins = [(11, 900), (1500, 1800), (8000, L)]
        g = ia.ScanWindows("cyk", cm, 1, dsq, wins, 0.0)
    finally:
        ia.unpin_target()
    u = ia.ScanWindows("cyk", cm, 1, dsq, wins, 0.0)
    assert hits_equal(sorted(g["hits"]), sorted(u["hits"]))
    assert np.float32(g["sc"]).view(np.uint32) == np.float32(u["sc"]).view(np.uint32)
    _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, 1, dsq, 2301, 6000, 0.0, env=3.0)


@pytest.mark.parametrize("name,qdb,L", [("rnaseP-eubact.c.cm", 1, 5000), ("Plant_SRP.c.cm", 1, 5000),
                                        ("synth_nested1500.cm", 2, 1800)])
def test_bifurcation_split_at_deck_ring_edge(name, qdb, L):
    """Band sets in which a bifurcation's largest split k is 63.  A state
    visit writes the BEGL deck rows jp .. jp+NR-1 before the bifurcation
    reads rows jp+t-k down to jp-63, so the deck ring must hold 63 + NR rows;
    a ring of 64 would hand the k = 63 term a row written in the same visit.
    CYK and FInside against the oracle, per-state maxima included."""
    cm = load_cm(name)
    dsq = random_dsq(L, seed=31)
    if qdb == 1:
        _run(ia.FastCYKScan, ob.ORACLE_CYK, cm, qdb, dsq, 1, L, 0.0, env=3.0, vsc=True)
    _run(ia.FastFInsideScan, ob.ORACLE_FINSIDE, cm, qdb, dsq, 1, L, 0.0, env=3.0, vsc=True)
