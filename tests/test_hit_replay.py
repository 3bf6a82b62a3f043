"""Host replay of the scanners' hit reporting (infx_resolve_hits, greedy
path) against a direct restatement of the reference on random candidate
sets, including ties in score and start:

  ReportHitsGreedily   src/cm_mx.c:7547-7628 (per j, report a cell when its
                       score beats every score reported at that j so far and
                       reaches the cutoff)
  overlap removal      src/cm_dpsearch.c:651-661: stable sort by score (desc)
                       then start (asc) (src/cm_tophits.c:220-245, qsort is a
                       merge sort in glibc), then the quadratic scan of
                       remove_or_mark_overlaps_one_seq_memeff
                       (src/cm_tophits.c:1097-1150): a hit is removed if it
                       shares a position with an earlier surviving hit.
"""
import ctypes as C
import os

import numpy as np
import pytest

import infernal_amd as ia

CMDIR = os.path.join(os.path.dirname(__file__), "golden", "cm")


class CAND(C.Structure):
    _fields_ = [("j", C.c_int32), ("d", C.c_int32), ("sc", C.c_float), ("r", C.c_int32), ("mode", C.c_int32)]


def reference(cands, cutoff):
    """cands: list of (j, d, sc, r) in (j, d) order."""
    hits = []
    by_j = {}
    for c in cands:
        by_j.setdefault(c[0], []).append(c)
    for j in sorted(by_j):
        mx = np.float32(-1e36)
        for (jj, d, sc, r) in by_j[j]:
            sc = np.float32(sc)
            if sc > mx and sc >= np.float32(cutoff) and sc > np.float32(-9.999e35):
                hits.append((jj - d + 1, jj, sc, r))
                mx = sc
    order = sorted(range(len(hits)), key=lambda k: (-hits[k][2], hits[k][0]))   # Python's sort is stable
    srt = [hits[k] for k in order]
    removed = [False] * len(srt)
    for a in range(len(srt)):
        if removed[a]:
            continue
        for b in range(a + 1, len(srt)):
            if removed[b]:
                continue
            if not (srt[b][1] < srt[a][0]) and not (srt[a][1] < srt[b][0]):
                removed[b] = True
    return [(h[0], h[1], float(h[2]), h[3]) for h, rm in zip(srt, removed) if not rm]


@pytest.mark.parametrize("seed", range(6))
def test_greedy_replay_matches_reference_restatement(seed):
    lib = ia.lib()
    f = lib.infx_resolve_hits
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                  C.c_int, C.c_float, C.c_int, C.c_int, C.c_void_p]
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    rng = np.random.default_rng(seed)
    i0, j0, W = 1, 3000, 60
    n = 4000
    js = rng.integers(i0, j0 + 1, n)
    ds = np.minimum(rng.integers(1, W + 1, n), js - i0 + 1)
    # coarse scores make ties in score (and start) common
    scs = (rng.integers(-20, 40, n) / 2.0).astype(np.float32)
    rs = rng.integers(0, 50, n)
    order = np.lexsort((ds, js))
    cands = [(int(js[k]), int(ds[k]), float(scs[k]), int(rs[k])) for k in order]
    # one candidate per (j, d): the scanners never repeat a cell
    seen, uniq = set(), []
    for c in cands:
        if (c[0], c[1]) not in seen:
            seen.add((c[0], c[1])); uniq.append(c)
    arr = (CAND * len(uniq))()
    perm = rng.permutation(len(uniq))       # the device hands them over in any order
    for a, k in enumerate(perm):
        j, d, sc, r = uniq[k]
        arr[a].j, arr[a].d, arr[a].sc, arr[a].r, arr[a].mode = j, d, sc, r, 3
    th = lib.cm_tophits_Create()
    eb = C.create_string_buffer(128)
    cutoff = 0.0
    st = f(C.cast(cm.ptr, C.c_void_p), eb, 0, 0, C.cast(arr, C.c_void_p), len(uniq), None, i0, j0, W, cutoff, 0, 1,
           C.cast(th, C.c_void_p))
    assert st == 0, eb.value
    got = [(h[0], h[1], h[4], h[2]) for h in ia._hits_list(th)]   # (start, stop, root, mode, score)
    lib.cm_tophits_Destroy(th)
    want = reference(uniq, cutoff)
    assert got == want
