"""The CM-only search pipeline around the GPU scanners (cmsearch --nohmm):
a CYK filter over the whole sequence, then Inside over the surviving
windows, with E-values.

  cyk_filter_windows  pli_cyk_seq_filter (src/cm_pipeline.c:3505-3618):
                      FastCYKScan with the tight bands (QDB1) and the
                      cutoff whose local-CYK P-value is F6 (default 1e-4);
                      each hit start..stop becomes the window
                      max(1, stop-(W-1)) .. min(L, start+(W-1)), hits in
                      position order, overlapping windows merged
  final_stage         pli_final_stage (src/cm_pipeline.c:3647-3812): in each
                      window FastIInsideScan with the loose bands (QDB2) and
                      the reporting threshold T = the score of E-value E
                      (E2ScoreGivenExpInfo, src/cm_pipeline.c:1069-1076);
                      P-values from the local-Inside tail
  search              both strands (the reverse complement searched as its
                      own sequence, its hits reported with start > stop),
                      E-values over the search space Z (default: both
                      strands of the sequence), hits with E <= E kept.

The GPU scanners take a whole sequence per call and split it internally
(chunks overlapping by W-1), so there is no 100 kb window reader.  Not
covered: the HMM filter stages, truncated-hit passes at sequence ends, and
hit alignment.
"""
import math

import numpy as np

from . import CM, FastCYKScan, ScanWindows, dsq_from_codes
from .calibrate import read_exp_infos


def cyk_filter_windows(cm, dsq, n, exp_lc, F6=1e-4, null3=True):
    """Windows (1-based, inclusive) that survive the CYK filter."""
    cutoff = exp_lc.mu_extrap + math.log(F6) / (-exp_lc.lambda_)
    r = FastCYKScan(cm, 1, dsq, 1, n, cutoff, do_null3=null3, want_hits=True)
    hits = sorted(r["hits"], key=lambda h: (h[0], h[1]))
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


def final_stage(cm, dsq, windows, exp_li, T, null3=True):
    """Inside hits (start, stop, score, P-value) in the windows: one
    FastIInsideScan launch for all of them (infx_FastScanWindows; each
    window's hits are those of one FastIInsideScan call on it)."""
    if not windows:
        return []
    r = ScanWindows("iinside", cm, 2, dsq, windows, T, do_null3=null3)
    hits = sorted(r["hits"], key=lambda h: (h[0], h[1]))
    return [(h[0], h[1], float(h[4]), float(exp_li.pvalue(h[4]))) for h in hits]


# complement of each residue code of the RNA alphabet "ACGU-RYMKSWHBVDN*~":
# A<->U, C<->G, R<->Y, M<->K, H<->D, B<->V; gap, S, W, N, '*' and '~' map to
# themselves (esl_sq_ReverseComplement with the alphabet's complement table)
_COMPLEMENT = np.array([3, 2, 1, 0, 4, 6, 5, 8, 7, 9, 10, 14, 13, 12, 11, 15, 16, 17], dtype=np.uint8)


def revcomp(codes):
    """Reverse complement of residue codes (canonical and IUPAC degenerate)."""
    return _COMPLEMENT[np.asarray(codes, dtype=np.uint8)[::-1]]


def search(cm_path, codes, Z=None, E=10.0, F6=1e-4, both_strands=True, null3=True, cm=None):
    """cmsearch --nohmm on one sequence of residue codes.  Returns
    (hits, stats): hits as dicts sorted by E-value, stats with the window
    counts per strand."""
    if cm is None:
        cm = CM.read(cm_path)
        cm.configure(local=True)
    exps = read_exp_infos(cm_path)
    if not exps:
        raise ValueError("model is not calibrated (no ECM lines)")
    codes = np.asarray(codes, dtype=np.uint8)
    n = len(codes)
    if Z is None:
        Z = float(n * (2 if both_strands else 1))
    lc, li = exps["LC"].update_for_dbsize(Z), exps["LI"].update_for_dbsize(Z)
    T = li.score_for_evalue(E)
    hits, stats = [], {}
    strands = [("+", codes)] + ([("-", revcomp(codes))] if both_strands else [])
    for strand, seq in strands:
        dsq = dsq_from_codes(seq)
        wins = cyk_filter_windows(cm, dsq, n, lc, F6, null3)
        found = final_stage(cm, dsq, wins, li, T, null3)
        stats[strand] = {"windows": len(wins), "residues": sum(j - i + 1 for i, j in wins), "hits": len(found)}
        for (s, e, sc, p) in found:
            if strand == "-":
                s, e = n - s + 1, n - e + 1
            hits.append({"start": s, "stop": e, "strand": strand, "score": sc, "pvalue": p, "evalue": p * li.cur_eff_dbsize})
    hits = [h for h in hits if h["evalue"] <= E]
    hits.sort(key=lambda h: (h["evalue"], -h["score"]))
    return hits, stats
