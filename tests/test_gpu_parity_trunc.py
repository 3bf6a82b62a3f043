"""GPU truncated scanners vs the oracle, through the C ABI.

RefTrCYKScan bit-exact and RefITrInsideScan exact (integer) in every truncated
pass (5' only, 3' only, 5' and 3' forced, 5' and 3' any), local and glocal:
scores, best mode, hit lists (with modes), envelopes and per-state maxima.
"""
import ctypes as C

import numpy as np
import pytest

import infernal_amd as ia
import oracle_bind as ob
from helpers import load_cm, cm_path, random_dsq, hits_equal, describe_hit_diff

pytestmark = pytest.mark.gpu

PASSES = [ia.PLI_PASS_5P_ONLY_FORCE, ia.PLI_PASS_3P_ONLY_FORCE, ia.PLI_PASS_5P_AND_3P_FORCE, ia.PLI_PASS_5P_AND_3P_ANY]


def _cmp(fn, algo, cm, qdb, p, dsq, i0, j0, cutoff, vsc=True, null3=False, greedy=True):
    print("case", fn.__name__, cm.c.name.decode(), qdb, p, i0, j0, cutoff, null3, greedy, flush=True)
    g = fn(cm, qdb, p, dsq, i0, j0, cutoff, do_null3=null3, env_cutoff=2.0, want_vsc=vsc, greedy=greedy)
    o = ob.scan(cm, algo, qdb, dsq, i0, j0, cutoff, do_null3=null3, greedy=greedy, env_cutoff=2.0, pass_idx=p)
    assert np.float32(g["sc"]).view(np.uint32) == np.float32(o["sc"]).view(np.uint32), (g["sc"], o["sc"])
    assert g["mode"] == o["mode"], (g["mode"], o["mode"])
    assert hits_equal(g["hits"], o["hits"]), describe_hit_diff(g["hits"], o["hits"])
    assert (g["envi"], g["envj"]) == (o["envi"], o["envj"])
    if vsc:
        assert np.array_equal(g["vsc"].view(np.uint32), o["vsc"].view(np.uint32)), np.nonzero(g["vsc"] != o["vsc"])


@pytest.mark.parametrize("p", PASSES)
@pytest.mark.parametrize("local", [True, False], ids=["local", "glocal"])
def test_trcyk_passes(p, local):
    cm = load_cm("tRNA.c.cm", local=local, trunc=True)
    dsq = random_dsq(400, seed=70 + p)
    _cmp(ia.RefTrCYKScan, ob.ORACLE_TRCYK, cm, 1, p, dsq, 1, 400, -10.0)


@pytest.mark.parametrize("p", PASSES)
def test_triinside_passes(p):
    cm = load_cm("tRNA.c.cm", local=True, trunc=True)
    dsq = random_dsq(350, seed=80 + p)
    _cmp(ia.RefITrInsideScan, ob.ORACLE_TRIINSIDE, cm, 2, p, dsq, 1, 350, -10.0)


def test_trunc_bifurcating_and_options():
    cm = load_cm("Vault.c.cm", local=True, trunc=True)
    dsq = random_dsq(500, seed=91)
    _cmp(ia.RefTrCYKScan, ob.ORACLE_TRCYK, cm, 1, ia.PLI_PASS_5P_AND_3P_FORCE, dsq, 21, 480, -5.0, null3=True)
    _cmp(ia.RefTrCYKScan, ob.ORACLE_TRCYK, cm, 2, ia.PLI_PASS_5P_AND_3P_ANY, dsq, 1, 500, 0.0, greedy=False)
    _cmp(ia.RefITrInsideScan, ob.ORACLE_TRIINSIDE, cm, 0, ia.PLI_PASS_3P_ONLY_FORCE, dsq, 1, 300, -5.0)
    cm2 = load_cm("Cobalamin.c.cm", local=True, trunc=True)
    dsq2 = random_dsq(600, seed=92)
    _cmp(ia.RefTrCYKScan, ob.ORACLE_TRCYK, cm2, 1, ia.PLI_PASS_5P_AND_3P_FORCE, dsq2, 1, 600, -5.0)
    _cmp(ia.RefITrInsideScan, ob.ORACLE_TRIINSIDE, cm2, 1, ia.PLI_PASS_5P_ONLY_FORCE, dsq2, 1, 600, -5.0)


def test_trunc_window_batch():
    """infx_TrScanWindows: many windows in one launch == one call per window."""
    cm = load_cm("tRNA.c.cm", local=True, trunc=True)
    dsq = random_dsq(3000, seed=93)
    rng = np.random.default_rng(5)
    n = 37
    i0 = rng.integers(1, 2500, n).astype(np.int64)
    j0 = (i0 + rng.integers(0, 450, n)).astype(np.int64)
    lib = ia.lib()
    lib.infx_TrScanWindows.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                       C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int, C.c_float,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    ths = [lib.cm_tophits_Create() for _ in range(n)]
    arr = (C.c_void_p * n)(*[C.cast(t, C.c_void_p) for t in ths])
    sc = np.zeros(n, np.float32); md = np.zeros(n, np.int8); ei = np.zeros(n, np.int64); ej = np.zeros(n, np.int64)
    cm.c.search_opts &= ~ia.CM_SEARCH_INSIDE
    eb = C.create_string_buffer(128)
    st = lib.infx_TrScanWindows(C.cast(cm.ptr, C.c_void_p), eb, cm.c.trsmx, 0, 1, ia.PLI_PASS_5P_AND_3P_FORCE,
                                dsq.ctypes.data, n, i0.ctypes.data, j0.ctypes.data, -5.0, arr, 0, 2.0,
                                sc.ctypes.data, md.ctypes.data, ei.ctypes.data, ej.ctypes.data)
    assert st == 0, eb.value
    for w in range(n):
        o = ob.scan(cm, ob.ORACLE_TRCYK, 1, dsq, int(i0[w]), int(j0[w]), -5.0, env_cutoff=2.0, pass_idx=ia.PLI_PASS_5P_AND_3P_FORCE)
        got = ia._hits_list(ths[w])
        assert np.float32(sc[w]).view(np.uint32) == np.float32(o["sc"]).view(np.uint32), (w, sc[w], o["sc"])
        assert md[w] == o["mode"]
        assert hits_equal(got, o["hits"]), (w, describe_hit_diff(got, o["hits"]))
        assert (ei[w], ej[w]) == (o["envi"], o["envj"])
        lib.cm_tophits_Destroy(ths[w])


def test_trunc_nested1500_clipped_windows():
    """The truncated-scan workload's shape: the 1,500-nt nested-stem model
    (W 1522) over windows of 200-320 nt, two holding an emitted homolog
    clipped by the window, two random decoys.  RefTrCYKScan in the 5'-only
    pass (QDB1) and RefITrInsideScan in the forced 5'-and-3' pass (QDB2), all
    windows in one launch each, against the oracle per window (run in threads:
    the oracle's truncated Inside takes ~10 s per window)."""
    import concurrent.futures as cf
    from infernal_amd import synth
    cm = load_cm("synth_nested1500.cm", local=True, trunc=True)
    codes, wins, hom = synth.truncated_windows(ia.CM.read(cm_path("synth_nested1500.cm")), 4, seed=301,
                                               lmin=200, lmax=320, pool=2)
    assert list(hom) == [True, False, True, False]
    dsq = ia.dsq_from_codes(codes)
    for algo, oalgo, q, p in (("cyk", ob.ORACLE_TRCYK, 1, ia.PLI_PASS_5P_ONLY_FORCE),
                              ("iinside", ob.ORACLE_TRIINSIDE, 2, ia.PLI_PASS_5P_AND_3P_FORCE)):
        print("case nested1500", algo, q, p, wins, flush=True)
        g = ia.TrScanWindows(algo, cm, q, p, dsq, wins, -5.0, env_cutoff=2.0)
        with cf.ThreadPoolExecutor(len(wins)) as ex:
            os_ = list(ex.map(lambda w: ob.scan(cm, oalgo, q, dsq, w[0], w[1], -5.0, env_cutoff=2.0, pass_idx=p), wins))
        for w, o in enumerate(os_):
            assert np.float32(g["sc"][w]).view(np.uint32) == np.float32(o["sc"]).view(np.uint32), (algo, w, g["sc"][w], o["sc"])
            assert g["mode"][w] == o["mode"], (algo, w)
            assert hits_equal(g["hits"][w], o["hits"]), (algo, w, describe_hit_diff(g["hits"][w], o["hits"]))
            assert (g["envi"][w], g["envj"][w]) == (o["envi"], o["envj"]), (algo, w)
        # the clipped homologs (one keeps most of the model, one a short
        # clipped piece) score above the decoys
        assert min(g["sc"][0], g["sc"][2]) > max(g["sc"][1], g["sc"][3]) + 10.0, g["sc"]
