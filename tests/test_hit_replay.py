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


def reference_fast(js, ds, scs, rs, cutoff, i0, j0):
    """The same restatement for large sets: candidates in (j, d) order as
    arrays; overlap removal over a boolean array of used positions (a hit is
    dropped if it shares a position with an earlier kept one, which is what
    the quadratic scan decides)."""
    hits = []
    n = len(js)
    k = 0
    while k < n:
        j = js[k]
        mx = np.float32(-1e36)
        while k < n and js[k] == j:
            sc = scs[k]
            if sc > mx and sc >= np.float32(cutoff) and sc > np.float32(-9.999e35):
                hits.append((int(j - ds[k] + 1), int(j), sc, int(rs[k])))
                mx = sc
            k += 1
    order = sorted(range(len(hits)), key=lambda q: (-hits[q][2], hits[q][0]))   # stable
    used = np.zeros(j0 - i0 + 2, dtype=bool)
    out = []
    for q in order:
        st, sp, sc, r = hits[q]
        if used[st - i0:sp - i0 + 1].any():
            continue
        used[st - i0:sp - i0 + 1] = True
        out.append((st, sp, float(sc), r))
    return out


@pytest.mark.parametrize("seed", range(2))
def test_greedy_replay_large_set(seed):
    """Candidate sets above the size where the replay orders candidates and
    hits by radix sorts (65,536) instead of comparison sorts: the same hits,
    in the same order, as the restatement; ties in score and start common,
    candidates handed over in a random order."""
    lib = ia.lib()
    f = lib.infx_resolve_hits
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                  C.c_int, C.c_float, C.c_int, C.c_int, C.c_void_p]
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    rng = np.random.default_rng(100 + seed)
    i0, j0, W = 5001, 405000, 60
    n = 300000
    js = rng.integers(i0, j0 + 1, n)
    ds = np.minimum(rng.integers(1, W + 1, n), js - i0 + 1)
    key = js.astype(np.int64) * 1024 + ds
    _, first = np.unique(key, return_index=True)       # one candidate per (j, d)
    js, ds = js[first], ds[first]
    m = len(js)
    assert m > 65536 * 3
    scs = (rng.integers(-20, 60, m) / 2.0).astype(np.float32)
    scs[rng.random(m) < 0.01] = np.float32(-0.0)        # -0 and +0 tie, as the comparison does
    rs = rng.integers(0, 50, m).astype(np.int32)
    perm = rng.permutation(m)
    arr = np.zeros(m, dtype=[("j", "<i4"), ("d", "<i4"), ("sc", "<f4"), ("r", "<i4"), ("mode", "<i4")])
    arr["j"], arr["d"], arr["sc"], arr["r"], arr["mode"] = js[perm], ds[perm], scs[perm], rs[perm], 3
    th = lib.cm_tophits_Create()
    eb = C.create_string_buffer(128)
    cutoff = 0.0
    st = f(C.cast(cm.ptr, C.c_void_p), eb, 0, 0, arr.ctypes.data, m, None, i0, j0, W, cutoff, 0, 1, C.cast(th, C.c_void_p))
    assert st == 0, eb.value
    got = [(h[0], h[1], h[4], h[2]) for h in ia._hits_list(th)]
    lib.cm_tophits_Destroy(th)
    o = np.lexsort((ds, js))
    want = reference_fast(js[o], ds[o], scs[o], rs[o], cutoff, i0, j0)
    assert len(got) > 1000
    assert got == want


def _allow(cm, pass_any, local, start, stop, i0, j0, mode, b):
    """cm_hit_AllowTruncation (src/cm_tophits.c:3048) from the begin state's
    node span, as the truncated kernel's filter evaluates it."""
    c = cm.c
    nd = c.ndidx[b]
    t = ord(c.ndtype[nd])
    MATP, MATL, MATR = 1, 2, 3     # node types (src/infernal.h)
    lpos = c.emap.contents.lpos[nd] if t in (MATP, MATL) else c.emap.contents.lpos[nd] + 1
    rpos = c.emap.contents.rpos[nd] if t in (MATP, MATR) else c.emap.contents.rpos[nd] - 1
    if pass_any or (start == i0 and stop == j0):
        return True
    if local:
        return mode == 3 or (mode == 2 and stop == j0) or (mode == 1 and start == i0)
    return (mode == 3 and lpos == 1 and rpos == c.clen) or (mode == 2 and stop == j0 and lpos == 1) or \
        (mode == 1 and start == i0 and rpos == c.clen)


@pytest.mark.parametrize("local", [True, False])
@pytest.mark.parametrize("pass_idx", [2, 3, 4])
def test_truncated_block_record_filter_keeps_every_reported_hit(local, pass_idx):
    """The truncated kernel, in greedy mode without NULL3, hands the host only
    the cells (j, d) that are admitted (cm_hit_AllowTruncation for their best
    mode and begin state), reach the cutoff and beat every admitted cell of
    a smaller d in the same 64-cell block of d.  ReportHitsGreedily
    (src/cm_mx.c:7547) reports a cell only when it beats every admitted
    smaller d of the row, so the dropped cells are never reported: the
    library's replay gives the same hits from the filtered set as from all
    cells.  Random rows with coarse scores (ties), all four modes and random
    begin states, on a window whose first and last rows are in the set."""
    lib = ia.lib()
    f = lib.infx_resolve_hits
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                  C.c_int, C.c_float, C.c_int, C.c_int, C.c_void_p]
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=local, trunc=True, trscanmx=True)
    rng = np.random.default_rng(100 * pass_idx + local)
    i0, j0, W = 1001, 1400, 150
    pass_any = pass_idx in (1, 5, 6)
    cells = []
    for j in range(i0, j0 + 1):
        dmax = min(W, j - i0 + 1)
        for d in range(1, dmax + 1):
            if rng.random() < 0.6:
                cells.append((j, d, float(rng.integers(-10, 30) / 2.0), int(rng.integers(1, cm.M - 1)), int(rng.integers(0, 4))))
    cells = [c for c in cells if ord(cm.c.sttype[c[3]]) != ia.E_st]
    kept = []
    by_row = {}
    for c in cells:
        by_row.setdefault(c[0], []).append(c)
    cutoff = 0.0
    for j, row in by_row.items():
        row.sort(key=lambda c: c[1])
        blk_max = {}
        for (jj, d, sc, r, mode) in row:
            b = (d - 1) // 64
            q = sc >= cutoff
            adm = q and _allow(cm, pass_any, local, jj - d + 1, jj, i0, j0, mode, r)
            prev = blk_max.get(b, -np.inf)
            if adm and sc > prev:
                kept.append((jj, d, sc, r, mode))
            if adm:
                blk_max[b] = max(prev, sc)

    def replay(cs):
        arr = (CAND * len(cs))()
        for a, k in enumerate(rng.permutation(len(cs))):
            j, d, sc, r, mode = cs[k]
            arr[a].j, arr[a].d, arr[a].sc, arr[a].r, arr[a].mode = j, d, sc, r, mode
        th = lib.cm_tophits_Create()
        eb = C.create_string_buffer(128)
        st = f(C.cast(cm.ptr, C.c_void_p), eb, pass_idx, 1, C.cast(arr, C.c_void_p), len(cs), None, i0, j0, W, cutoff, 0, 1,
               C.cast(th, C.c_void_p))
        assert st == 0, eb.value
        out = sorted(ia._hits_list(th))
        lib.cm_tophits_Destroy(th)
        return out

    everything, filtered = replay(cells), replay(kept)
    assert len(kept) < len(cells) // 4
    assert len(everything) > 0
    assert filtered == everything


def test_replay_refuses_candidates_outside_the_range():
    """A candidate whose cell lies outside the call's range (start before i0,
    end past j0, d < 1) is refused with eslEINVAL instead of being indexed."""
    lib = ia.lib()
    f = lib.infx_resolve_hits
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                  C.c_int, C.c_float, C.c_int, C.c_int, C.c_void_p]
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    i0, j0, W = 101, 400, 60
    for j, d in ((120, 25), (401, 3), (200, 0), (150, 40)):
        arr = (CAND * 2)()
        arr[0].j, arr[0].d, arr[0].sc, arr[0].r, arr[0].mode = 300, 10, 5.0, 0, 3
        arr[1].j, arr[1].d, arr[1].sc, arr[1].r, arr[1].mode = j, d, 5.0, 0, 3
        th = lib.cm_tophits_Create()
        eb = C.create_string_buffer(128)
        st = f(C.cast(cm.ptr, C.c_void_p), eb, 0, 0, C.cast(arr, C.c_void_p), 2, None, i0, j0, W, 0.0, 0, 1, C.cast(th, C.c_void_p))
        lib.cm_tophits_Destroy(th)
        ok = (j - d + 1 >= i0) and j <= j0 and d >= 1
        assert (st == 0) == ok, (j, d, st, eb.value)
        if not ok:
            assert st == ia.eslEINVAL and b"outside" in eb.value
