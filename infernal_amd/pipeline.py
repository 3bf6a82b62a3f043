"""The CM-only search pipeline around the GPU scanners (cmsearch --nohmm):
a CYK filter over each sequence window, then Inside over the surviving
windows, with E-values; truncated passes at the sequence ends.

  read_windows        esl_sqio_ReadWindow as cmsearch uses it
                      (src/cmsearch.c:802-879): windows of 100,000 new
                      residues (CM_MAX_RESIDUE_COUNT, src/infernal.h:116),
                      each after the first preceded by maxW residues of the
                      previous one, maxW = max(W, 1.25 clen)
                      (src/cm_pipeline.c:1048-1055)
  pipeline_passes     the passes cm_Pipeline runs on a window
                      (src/cm_pipeline.c:1325-1412, 1430-1471, default
                      truncation): the standard pass over the window; with
                      the sequence's 5' end in the window, a 5'-truncation
                      pass over its first maxW residues; with the 3' end, a
                      3'-truncation pass over its last maxW; with both and
                      at most maxW residues, a 5'-and-3' pass over it all
  cyk_filter_windows  pli_cyk_seq_filter (src/cm_pipeline.c:3505-3618):
                      FastCYKScan (truncated passes: RefTrCYKScan) with the
                      tight bands (QDB1) and the cutoff whose local-CYK
                      P-value is F6 (default 1e-4); each hit start..stop
                      becomes the window max(1, stop-(W-1)) ..
                      min(L, start+(W-1)), hits in position order,
                      overlapping windows merged
  final_stage         pli_final_stage (src/cm_pipeline.c:3647-3812): in each
                      window FastIInsideScan (truncated passes:
                      RefITrInsideScan) with the loose bands (QDB2) and the
                      reporting threshold T = the score of E-value E
                      (E2ScoreGivenExpInfo, src/cm_pipeline.c:1069-1076);
                      P-values from the local-Inside tail
  search              both strands (each window's reverse complement searched
                      as its own sequence), hit coordinates mapped to the
                      sequence (3'-pass hits shifted by their offset,
                      src/cm_pipeline.c:1699-1712; window offsets,
                      cm_tophits_UpdateHitPositions), overlapping hits of one
                      strand removed best first (cm_tophits_RemoveOrMarkOverlaps
                      after the search, src/cmsearch.c:665), E-values over
                      the search space Z (default: both strands), hits with
                      E <= E kept.

Not covered: the HMM filter stages and hit alignment.
"""
import bisect
import math

import numpy as np

from . import (CM, FastCYKScan, ScanWindows, TrScanWindows, dsq_from_codes, PLI_PASS_STD_ANY, PLI_PASS_5P_ONLY_FORCE,
               PLI_PASS_3P_ONLY_FORCE, PLI_PASS_5P_AND_3P_FORCE)
from .calibrate import read_exp_infos


CM_MAX_RESIDUE_COUNT = 100000   # residues per sequence window (src/infernal.h:116)


def max_w(cm):
    """pli->maxW: max(W, 1.25 clen), truncated to an int (src/cm_pipeline.c:1055)."""
    return int(max(1.0 * cm.W, 1.25 * cm.clen))


def read_windows(L, maxW, window=CM_MAX_RESIDUE_COUNT):
    """(start, end) of the windows esl_sqio_ReadWindow(maxW, window) returns
    for a sequence of L residues: the first holds residues 1..window, each
    later one maxW residues of context before the next `window` new ones."""
    out = []
    a, b = 1, min(L, window)
    out.append((a, b))
    while b < L:
        a, b = max(1, b - maxW + 1), min(L, b + window)
        out.append((a, b))
    return out


def pipeline_passes(n, have5, have3, maxW, trunc=True):
    """The passes of cm_Pipeline on an n-residue window (default truncation
    options): (pass index, first, last residue searched)."""
    out = [(PLI_PASS_STD_ANY, 1, n)]
    if trunc:
        if have5:
            out.append((PLI_PASS_5P_ONLY_FORCE, 1, min(n, maxW)))
        if have3:
            out.append((PLI_PASS_3P_ONLY_FORCE, n - min(n, maxW) + 1, n))
        if have5 and have3 and n <= maxW:
            out.append((PLI_PASS_5P_AND_3P_FORCE, 1, n))
    return out


def cyk_filter_windows(cm, dsq, n, exp_lc, F6=1e-4, null3=True, pass_idx=PLI_PASS_STD_ANY):
    """Windows (1-based, inclusive) that survive the CYK filter."""
    cutoff = exp_lc.mu_extrap + math.log(F6) / (-exp_lc.lambda_)
    if pass_idx == PLI_PASS_STD_ANY:
        r = FastCYKScan(cm, 1, dsq, 1, n, cutoff, do_null3=null3, want_hits=True)
        found = r["hits"]
    else:
        found = TrScanWindows("cyk", cm, 1, pass_idx, dsq, [(1, n)], cutoff, do_null3=null3)["hits"][0]
    hits = sorted(found, key=lambda h: (h[0], h[1]))
    W = cm.W
    wins = []
    h = 0
    while h < len(hits):
        iwin = max(1, hits[h][1] - (W - 1))
        jwin = min(n, hits[h][0] + (W - 1))
        while h + 1 < len(hits) and max(1, hits[h + 1][1] - (W - 1)) <= jwin:
            h += 1
            jwin = min(n, hits[h][0] + (W - 1))
        wins.append((iwin, jwin))
        h += 1
    return wins


def final_stage(cm, dsq, windows, exp_li, T, null3=True, pass_idx=PLI_PASS_STD_ANY):
    """Inside hits (start, stop, score, P-value) in the windows: one launch
    for all of them, each window's hits those of one FastIInsideScan (or,
    in a truncated pass, RefITrInsideScan) call on it."""
    if not windows:
        return []
    if pass_idx == PLI_PASS_STD_ANY:
        found = ScanWindows("iinside", cm, 2, dsq, windows, T, do_null3=null3)["hits"]
    else:
        r = TrScanWindows("iinside", cm, 2, pass_idx, dsq, windows, T, do_null3=null3)
        found = [h for hw in r["hits"] for h in hw]
    hits = sorted(found, key=lambda h: (h[0], h[1]))
    return [(h[0], h[1], float(h[4]), float(exp_li.pvalue(h[4]))) for h in hits]


# complement of each residue code of the RNA alphabet "ACGU-RYMKSWHBVDN*~":
# A<->U, C<->G, R<->Y, M<->K, H<->D, B<->V; gap, S, W, N, '*' and '~' map to
# themselves (esl_sq_ReverseComplement with the alphabet's complement table)
_COMPLEMENT = np.array([3, 2, 1, 0, 4, 6, 5, 8, 7, 9, 10, 14, 13, 12, 11, 15, 16, 17], dtype=np.uint8)


def revcomp(codes):
    """Reverse complement of residue codes (canonical and IUPAC degenerate)."""
    return _COMPLEMENT[np.asarray(codes, dtype=np.uint8)[::-1]]


def search(cm_path, codes, Z=None, E=10.0, F6=1e-4, both_strands=True, null3=True, cm=None, trunc=True,
           window=CM_MAX_RESIDUE_COUNT):
    """cmsearch --nohmm on one sequence of residue codes, read in windows of
    `window` new residues.  trunc: run the truncated passes at the
    sequence's ends (the reference's default); False as --notrunc.
    Returns (hits, stats): hits as dicts sorted by E-value, stats with the
    windows, passes and filter survivors per strand."""
    if cm is None:
        cm = CM.read(cm_path)
        cm.configure(local=True, trunc=trunc, scanmx=True, trscanmx=trunc)
    exps = read_exp_infos(cm_path)
    if not exps:
        raise ValueError("model is not calibrated (no ECM lines)")
    codes = np.asarray(codes, dtype=np.uint8)
    L = len(codes)
    if Z is None:
        Z = float(L * (2 if both_strands else 1))
    lc, li = exps["LC"].update_for_dbsize(Z), exps["LI"].update_for_dbsize(Z)
    T = li.score_for_evalue(E)
    maxW = max_w(cm)
    hits, stats = [], {}
    for strand in (["+", "-"] if both_strands else ["+"]):
        st = {"windows": 0, "passes": 0, "filter_windows": 0, "filter_residues": 0, "hits": 0}
        for (a, b) in read_windows(L, maxW, window):
            seq = codes[a - 1:b] if strand == "+" else revcomp(codes[a - 1:b])
            n = b - a + 1
            # the strand's 5' and 3' sequence ends (src/cm_pipeline.c:1334-1341)
            have5, have3 = (a == 1, b == L) if strand == "+" else (b == L, a == 1)
            st["windows"] += 1
            for (pidx, x0, x1) in pipeline_passes(n, have5, have3, maxW, trunc):
                sub = seq[x0 - 1:x1]
                dsq = dsq_from_codes(sub)
                m = x1 - x0 + 1
                wins = cyk_filter_windows(cm, dsq, m, lc, F6, null3, pass_idx=pidx)
                found = final_stage(cm, dsq, wins, li, T, null3, pass_idx=pidx)
                st["passes"] += 1
                st["filter_windows"] += len(wins)
                st["filter_residues"] += sum(j - i + 1 for i, j in wins)
                for (hs, he, sc, pv) in found:
                    hs, he = hs + x0 - 1, he + x0 - 1          # in the window (3'-pass offset)
                    if strand == "+":
                        gs, ge = a + hs - 1, a + he - 1
                    else:
                        gs, ge = b - hs + 1, b - he + 1
                    hits.append({"start": gs, "stop": ge, "strand": strand, "score": sc, "pvalue": float(pv),
                                 "evalue": pv * li.cur_eff_dbsize, "pass": pidx})
        stats[strand] = st
    hits = remove_overlaps(hits)
    for strand in stats:
        stats[strand]["hits"] = sum(1 for h in hits if h["strand"] == strand)
    hits = [h for h in hits if h["evalue"] <= E]
    hits.sort(key=lambda h: (h["evalue"], -h["score"]))
    return hits, stats


def remove_overlaps(hits):
    """Overlap removal over the whole search (cm_tophits_SortForOverlapRemoval
    + remove_or_mark_overlaps_one_seq_memeff, src/cm_tophits.c:220-245,
    1097-1150): per strand, by score descending then start ascending; a hit
    sharing a residue with a kept hit is dropped.  The kept intervals are
    disjoint, so they are held sorted by their left end and a hit is checked
    against the one kept interval that starts last at or before its right
    end (O(log n) per hit)."""
    out = []
    for strand in sorted({h["strand"] for h in hits}):
        hs = sorted((h for h in hits if h["strand"] == strand), key=lambda h: (-h["score"], h["start"]))
        los, his = [], []
        for h in hs:
            lo, hi = min(h["start"], h["stop"]), max(h["start"], h["stop"])
            k = bisect.bisect_right(los, hi)
            if k and his[k - 1] >= lo:
                continue
            los.insert(k, lo)
            his.insert(k, hi)
            out.append(h)
    return out