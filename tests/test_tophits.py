"""The exported hit-list functions a caller uses after the scanners, in the
manner of the reference's cm_tophits_utest (src/cm_tophits.c:3349-3795:
simulated hits, then sorting and overlap removal checked):

  cm_tophits_SortForOverlapRemoval  src/cm_tophits.c:400-420 with the
                                    comparator at :220-245: by model, sequence,
                                    strand, score (desc), start (asc); stable
  cm_tophits_RemoveOrMarkOverlaps   src/cm_tophits.c:1203-1275 with
                                    remove_or_mark_overlaps_one_seq_memeff
                                    (:1097-1150): within one model, sequence
                                    and strand, a hit that shares a position
                                    with an earlier kept hit is marked a
                                    removed duplicate (not reported, not
                                    included); hits already so marked are
                                    skipped
  cm_tophits_CloneHitMostly         src/cm_tophits.c:877-915

against a direct quadratic restatement, on random hits with ties in score
and start, both strands (minus-strand hits have start > stop), and hits
already marked as duplicates.
"""
import ctypes as C

import numpy as np
import pytest

import infernal_amd as ia

INCLUDED, REPORTED, DUP = 1, 2, 4


def _lib():
    lib = ia.lib()
    lib.cm_tophits_CreateNextHit.argtypes = [C.POINTER(ia.CM_TOPHITS), C.POINTER(C.POINTER(ia.CM_HIT))]
    lib.cm_tophits_SortForOverlapRemoval.argtypes = [C.POINTER(ia.CM_TOPHITS)]
    lib.cm_tophits_RemoveOrMarkOverlaps.argtypes = [C.POINTER(ia.CM_TOPHITS), C.c_int, C.c_char_p]
    lib.cm_tophits_CloneHitMostly.argtypes = [C.POINTER(ia.CM_TOPHITS), C.c_int, C.POINTER(ia.CM_TOPHITS)]
    return lib


def make_list(lib, hits):
    th = lib.cm_tophits_Create()
    for (cmi, seqi, rc, start, stop, sc, flags, srcL) in hits:
        hp = C.POINTER(ia.CM_HIT)()
        assert lib.cm_tophits_CreateNextHit(th, C.byref(hp)) == 0
        h = hp.contents
        h.cm_idx, h.seq_idx, h.in_rc = cmi, seqi, rc
        h.start, h.stop, h.score, h.flags, h.srcL = start, stop, sc, flags, srcL
    return th


def reference(hits):
    """Creation indices in the sorted order, and the set removed."""
    order = sorted(range(len(hits)), key=lambda k: (hits[k][0], hits[k][1], hits[k][2], -hits[k][5], hits[k][3]))
    removed = set()
    groups = {}
    for k in order:
        groups.setdefault(hits[k][:3], []).append(k)
    for g in groups.values():
        kept = []
        for k in g:
            if hits[k][6] & DUP:
                continue
            lo, hi = min(hits[k][3], hits[k][4]), max(hits[k][3], hits[k][4])
            if any(not (hi < klo or lo > khi) for (klo, khi) in kept):
                removed.add(k)
            else:
                kept.append((lo, hi))
    return order, removed


@pytest.mark.parametrize("seed", range(5))
def test_sort_and_overlap_removal(seed):
    lib = _lib()
    rng = np.random.default_rng(300 + seed)
    srcL = 2000
    hits = []
    for _ in range(600):
        cmi, seqi, rc = int(rng.integers(0, 2)), int(rng.integers(0, 3)), int(rng.integers(0, 2))
        a = int(rng.integers(1, srcL - 60))
        b = a + int(rng.integers(0, 60))
        start, stop = (b, a) if rc else (a, b)
        sc = float(rng.integers(0, 30)) / 2.0                # coarse: ties are common
        flags = DUP if rng.random() < 0.05 else (INCLUDED | REPORTED)
        hits.append((cmi, seqi, rc, start, stop, sc, flags, srcL))
    th = make_list(lib, hits)
    t = th.contents
    assert lib.cm_tophits_SortForOverlapRemoval(th) == 0
    # each sorted entry points into the creation-ordered array: its index is the creation index
    base = C.addressof(t.unsrt.contents)
    got_order = [(C.addressof(t.hit[i].contents) - base) // C.sizeof(ia.CM_HIT) for i in range(t.N)]
    eb = C.create_string_buffer(128)
    assert lib.cm_tophits_RemoveOrMarkOverlaps(th, 0, eb) == 0, eb.value
    got_removed = {k for k in range(t.N) if (t.unsrt[k].flags & DUP) and not (hits[k][6] & DUP)}
    for k in got_removed:
        assert not (t.unsrt[k].flags & (REPORTED | INCLUDED))
    want_order, want_removed = reference(hits)
    assert got_order == want_order
    assert got_removed == want_removed
    # a clone copies the hit but keeps its own index in the destination
    dst = lib.cm_tophits_Create()
    assert lib.cm_tophits_CloneHitMostly(th, 3, dst) == 0
    d = dst.contents.unsrt[0]
    s = t.hit[3].contents
    assert (d.start, d.stop, d.score, d.cm_idx, d.seq_idx, d.in_rc, d.hit_idx) == (s.start, s.stop, s.score, s.cm_idx, s.seq_idx, s.in_rc, 0)
    lib.cm_tophits_Destroy(dst)
    lib.cm_tophits_Destroy(th)


def test_overlap_removal_contract():
    """Unsorted lists, inconsistent source lengths and positions outside
    1..srcL are refused (src/cm_tophits.c:1218-1240)."""
    lib = _lib()
    eb = C.create_string_buffer(128)
    th = make_list(lib, [(0, 0, 0, 5, 9, 1.0, 0, 100), (0, 0, 0, 7, 12, 2.0, 0, 100)])
    assert lib.cm_tophits_RemoveOrMarkOverlaps(th, 0, eb) != 0            # not sorted
    lib.cm_tophits_Destroy(th)
    th = make_list(lib, [(0, 0, 0, 5, 9, 1.0, 0, 100), (0, 0, 0, 7, 12, 2.0, 0, 200)])
    assert lib.cm_tophits_SortForOverlapRemoval(th) == 0
    assert lib.cm_tophits_RemoveOrMarkOverlaps(th, 0, eb) != 0 and b"srcL" in eb.value
    lib.cm_tophits_Destroy(th)
    th = make_list(lib, [(0, 0, 0, 95, 101, 1.0, 0, 100), (0, 0, 0, 7, 12, 2.0, 0, 100)])
    assert lib.cm_tophits_SortForOverlapRemoval(th) == 0
    assert lib.cm_tophits_RemoveOrMarkOverlaps(th, 0, eb) != 0
    lib.cm_tophits_Destroy(th)
