"""Pins for the oracle's hit-reporting rules that no scan-level check pins:
the gamma (non-greedy) hit selection, truncated-hit admissibility and the
NULL3 score correction.

* Gamma (src/cm_mx.c:7440-7669, UpdateGammaHitMx / TBackGammaHitMx): the
  semi-HMM keeps the set of pairwise disjoint candidate hits with the
  largest total score.  Checked against exhaustive enumeration of every way
  to cover a short target with disjoint hits, on integer-valued scores (exact
  float sums), with impossible cells; the reported hits are the traced set
  filtered by the cutoff.
* cm_hit_AllowTruncation (src/cm_tophits.c:3048-3091): a truncated hit may
  lack model columns only where its marginal mode truncates the model and
  only at the window's end on that side; glocal hits must otherwise cover the
  model to its first and last consensus column, local ones may miss columns
  anywhere.  The consensus columns a state's sub-model covers are computed
  here by walking the node tree (left emissions on entry, right ones on exit),
  not from the emit map the library builds.  Every hit the truncated scanners
  report obeys the rule.
* ScoreCorrectionNull3 (src/cm_parsetree.c:2352-2380): closed forms.  A hit
  whose composition equals the null model gets log2(1 + 2^-16) (the 1/2^16
  prior alone), also when the composition comes from degenerate residues; a
  run of n copies of one residue of background probability p gets
  log2(1 + 2^(n log2(1/p) - 16)), i.e. log2(1 + 2^(2n-16)) under a uniform
  null.  The oracle sums through the reference's 0.001-bit log-sum table,
  so the bar is 1e-3 bits.
"""
import ctypes as C
import itertools
import math
import sys

import numpy as np
import pytest

import infernal_amd as ia
import oracle_bind as ob
from helpers import load_cm, random_dsq

IMP = np.float32(-1e36)


def _lib():
    lib = ob.load()
    lib.oracle_trunc_hits_begin.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_int]
    lib.oracle_trunc_hits_update.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_int]
    lib.oracle_trunc_hits_finish.argtypes = [C.c_void_p, C.POINTER(ob.ORACLE_RESULT), C.c_char_p]
    lib.oracle_allow_truncation.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int]
    lib.oracle_null3_hit.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64]
    lib.oracle_null3_hit.restype = C.c_float
    return lib


def gamma_select(cm, sc, i0, cutoff):
    """Drive the oracle's gamma reporter over rows sc[jp][d] (jp = 1..L,
    d = 0..W); returns the reported hits (start, stop, root, score)."""
    lib = _lib()
    L, W = sc.shape[0] - 1, sc.shape[1] - 1
    acc = C.c_void_p()
    lib.oracle_trunc_hits_begin(C.byref(acc), L, i0, i0 + L - 1, W, 0)
    for jp in range(1, L + 1):
        row = np.ascontiguousarray(sc[jp], dtype=np.float32)
        bestr = np.arange(W + 1, dtype=np.int32) + 100 * jp
        lib.oracle_trunc_hits_update(acc, C.cast(cm.ptr, C.c_void_p), ia.PLI_PASS_STD_ANY, i0 + jp - 1, 1, min(W, jp),
                                     row.ctypes.data, bestr.ctypes.data, None, None, cutoff, 0)
    res = ob.ORACLE_RESULT()
    eb = C.create_string_buffer(128)
    assert lib.oracle_trunc_hits_finish(acc, C.byref(res), eb) == 0
    out = [(res.hits[h].start, res.hits[h].stop, res.hits[h].root, res.hits[h].score) for h in range(res.nhits)]
    lib.oracle_result_free(C.byref(res))
    return out


def best_cover(sc, i0):
    """Exhaustive: every segmentation of 1..L into uncovered residues and
    hits of length <= W with a possible score; the best total score."""
    L, W = sc.shape[0] - 1, sc.shape[1] - 1
    best = 0.0
    stack = [(1, 0.0)]
    while stack:
        p, tot = stack.pop()
        if p > L:
            best = max(best, tot)
            continue
        stack.append((p + 1, tot))                      # residue p not in a hit
        for d in range(1, W + 1):                       # a hit p..p+d-1
            jp = p + d - 1
            if jp <= L and sc[jp][d] > -1e35:
                stack.append((jp + 1, tot + float(sc[jp][d])))
    return best


@pytest.mark.parametrize("seed", range(40))
def test_gamma_selects_the_best_disjoint_set(seed):
    cm = load_cm("tRNA.c.cm")
    rng = np.random.default_rng(seed)
    L, W, i0 = int(rng.integers(4, 11)), int(rng.integers(1, 5)), int(rng.integers(1, 50))
    sc = rng.integers(-6, 9, size=(L + 1, W + 1)).astype(np.float32)
    sc[rng.random(sc.shape) < 0.2] = IMP
    traced = gamma_select(cm, sc, i0, -1e30)
    # disjoint, each hit's own cell score, and the best total
    iv = sorted((h[0], h[1]) for h in traced)
    assert all(b[0] > a[1] for a, b in zip(iv, iv[1:]))
    for s, e, r, x in traced:
        jp, d = e - i0 + 1, e - s + 1
        assert x == sc[jp][d] and r == d + 100 * jp
    assert sum(h[3] for h in traced) == best_cover(sc, i0)
    # the cutoff only filters the traced set
    cut = 2.0
    assert gamma_select(cm, sc, i0, cut) == [h for h in traced if h[3] >= cut]


def cover_spans(cm):
    """First and last consensus column of each node's sub-model (a node
    without emissions below it: the span between its neighbours, first =
    last + 1)."""
    c = cm.c
    nodes = c.nodes
    ndt = [ord(c.ndtype[n]) if isinstance(c.ndtype[n], bytes) else c.ndtype[n] for n in range(nodes)]
    first, last = [0] * nodes, [0] * nodes
    col = 0
    stack = [("in", 0)]
    while stack:
        kind, nd = stack.pop()
        t = ndt[nd]
        if kind == "out":
            if t in (1, 3):            # MATP_nd, MATR_nd: right emission on the way out
                col += 1
            last[nd] = col
            continue
        first[nd] = col + 1
        if t in (1, 2):                # MATP_nd, MATL_nd: left emission on the way in
            col += 1
        stack.append(("out", nd))
        if t == 0:                     # BIF_nd: left sub-tree, then right
            b = c.nodemap[nd]
            stack.append(("in", c.ndidx[c.cnum[b]]))
            stack.append(("in", c.ndidx[c.cfirst[b]]))
        elif t != 7:                   # not END_nd: the next node
            stack.append(("in", nd + 1))
    assert col == c.clen
    return first, last


def admissible(local, clen, pass_idx, start, stop, i0, j0, mode, first, last):
    if pass_idx in (ia.PLI_PASS_STD_ANY, ia.PLI_PASS_5P_AND_3P_ANY, 6):   # 6: PLI_PASS_HMM_ONLY_ANY
        return True
    if start == i0 and stop == j0:
        return True
    cut5 = mode in (ia.TRMODE_R, ia.TRMODE_T)      # the model's 5' part is missing by mode
    cut3 = mode in (ia.TRMODE_L, ia.TRMODE_T)
    if mode == ia.TRMODE_T:
        return False
    if not local and ((first > 1 and not cut5) or (last < clen and not cut3)):
        return False                                # glocal: missing columns must be the mode's
    return (not cut5 or start == i0) and (not cut3 or stop == j0)


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Vault.c.cm", "Cobalamin.c.cm", "5S_rRNA.c.cm", "synth_nested1500.cm"])
@pytest.mark.parametrize("local", [True, False], ids=["local", "glocal"])
def test_allow_truncation_rule(name, local):
    cm = load_cm(name, local=local, trunc=True)
    lib = _lib()
    first, last = cover_spans(cm)
    c = cm.c
    i0, j0 = 101, 400
    for b in range(c.M):
        nd = c.ndidx[b]
        for mode in (ia.TRMODE_J, ia.TRMODE_L, ia.TRMODE_R, ia.TRMODE_T):
            for p in (1, 2, 3, 4, 5):
                for start, stop in ((i0, j0), (i0, j0 - 7), (i0 + 3, j0), (i0 + 3, j0 - 7)):
                    got = lib.oracle_allow_truncation(C.cast(cm.ptr, C.c_void_p), p, start, stop, i0, j0, mode, b)
                    want = admissible(local, c.clen, p, start, stop, i0, j0, mode, first[nd], last[nd])
                    assert bool(got) == want, (b, nd, mode, p, start, stop, first[nd], last[nd])
                    # the library's own function (what the GPU scanners' replay applies)
                    got_lib = ia.lib().cm_hit_AllowTruncation(cm.ptr, p, start, stop, i0, j0, bytes([mode]), b)
                    assert bool(got_lib) == want, ("library", b, nd, mode, p, start, stop)


@pytest.mark.parametrize("local", [True, False], ids=["local", "glocal"])
@pytest.mark.parametrize("p", [ia.PLI_PASS_5P_ONLY_FORCE, ia.PLI_PASS_3P_ONLY_FORCE, ia.PLI_PASS_5P_AND_3P_FORCE])
def test_reported_truncated_hits_are_admissible(local, p):
    cm = load_cm("tRNA.c.cm", local=local, trunc=True)
    first, last = cover_spans(cm)
    dsq = random_dsq(300, seed=500 + p)
    for algo, q in ((ob.ORACLE_TRCYK, 1), (ob.ORACLE_TRIINSIDE, 2)):
        o = ob.scan(cm, algo, q, dsq, 21, 280, -500.0, pass_idx=p)
        assert o["hits"]
        for s, e, root, mode, _ in o["hits"]:
            nd = cm.c.ndidx[root]
            assert admissible(local, cm.c.clen, p, s, e, 21, 280, mode, first[nd], last[nd])


def _null3(cm, codes):
    lib = _lib()
    dsq = ia.dsq_from_codes(np.asarray(codes, dtype=np.uint8))
    return float(lib.oracle_null3_hit(C.cast(cm.ptr, C.c_void_p), dsq.ctypes.data, 1, len(codes)))


def test_null3_closed_forms_uniform_null():
    cm = load_cm("tRNA.c.cm")
    assert np.allclose([cm.c.null[a] for a in range(4)], 0.25)
    prior = math.log2(1 + 2.0 ** -16)
    for k in (1, 3, 10, 40):
        assert abs(_null3(cm, [0, 1, 2, 3] * k) - prior) < 1e-3
        assert abs(_null3(cm, [15] * (4 * k)) - prior) < 1e-3           # N: a quarter of each residue
        assert abs(_null3(cm, [5, 6] * (2 * k)) - prior) < 1e-3         # R, Y: A/G and C/U halves
    for n in range(1, 31):
        want = math.log2(1 + 2.0 ** (2 * n - 16))
        for x in range(4):
            assert abs(_null3(cm, [x] * n) - want) < 1e-3, (n, x)


def test_null3_closed_forms_skewed_null():
    cm = ia.CM.read(ob.os.path.join(ob.ROOT, "tests", "golden", "cm", "tRNA.c.cm"))
    cm.configure(local=True)
    p = [0.1, 0.2, 0.3, 0.4]
    for a in range(4):
        cm.c.null[a] = p[a]
    prior = math.log2(1 + 2.0 ** -16)
    for k in (1, 2, 5):
        assert abs(_null3(cm, ([0] + [1] * 2 + [2] * 3 + [3] * 4) * k) - prior) < 1e-3
    for n in (1, 4, 9, 17):
        for x in range(4):
            want = math.log2(1 + 2.0 ** (n * math.log2(1 / p[x]) - 16))
            assert abs(_null3(cm, [x] * n) - want) < 1e-3, (n, x)
