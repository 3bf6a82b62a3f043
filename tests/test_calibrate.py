"""E-values and exponential-tail calibration (infernal_amd.calibrate).

Pinned against the reference's files and against closed forms:
* every calibrated model shipped with the reference (and every synthetic
  model calibrated here) satisfies
  mu_extrap = mu_orig - ln(1/tailp)/lambda (SetExpInfo) and
  tailp = (750 local | 250 glocal) x 1.6 Mb / nrandhits (fit_histogram);
* the E-values of all 56 hits of the tutorial's cmsearch run
  (tRNA5.c.cm vs mrum-genome.fa, 5,874,406 residues searched) follow from
  their scores, the model's ECMLI line and UpdateExpsForDBSize, within the
  output's rounding (scores to 0.1 bit, E-values to 2 digits);
* the tail fit recovers a planted exponential; score_for_evalue inverts
  evalue; the generative HMM's samples have its stationary composition.
"""
import math
import os

import numpy as np
import pytest

import infernal_amd as ia
from infernal_amd import calibrate as cal

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CMDIR = os.path.join(GOLD, "cm")
def _has_exp_lines(f):
    with open(os.path.join(CMDIR, f)) as fh:
        return any(line.startswith("ECMLC") for line in fh)


# the reference's calibrated models, and the synthetic benchmark models this
# library calibrated (tests/golden/calibrate_synthetic_cms.py)
CALIBRATED = [f for f in sorted(os.listdir(CMDIR)) if f.endswith(".cm") and _has_exp_lines(f)]


@pytest.mark.parametrize("name", CALIBRATED)
def test_shipped_exp_lines_closed_form(name):
    infos = cal.read_exp_infos(os.path.join(CMDIR, name))
    assert set(infos) == {"LC", "GC", "LI", "GI"}
    for mode, e in infos.items():
        mux = e.mu_orig - math.log(1.0 / e.tailp) / e.lambda_
        assert abs(mux - e.mu_extrap) < 2e-3, (mode, mux, e.mu_extrap)
        want_tp = cal.TAILN[mode[0]] * (e.dbsize / 1e6) / e.nrandhits
        assert abs(want_tp - e.tailp) <= 6e-7 + 1e-4 * want_tp, (mode, want_tp, e.tailp)


def _tutorial_hits():
    rows, z = [], None
    with open(os.path.join(GOLD, "tRNA5_mrum_hits.tsv")) as fh:
        for line in fh:
            if line.startswith("# cmsearch"):
                z = int(line.rsplit(None, 1)[1])
            elif not line.startswith("#"):
                f = line.split("\t")
                rows.append((float(f[1]), float(f[2])))
    return rows, z


def test_tutorial_evalues():
    rows, Z = _tutorial_hits()
    assert len(rows) == 56 and Z == 5874406
    e = cal.read_exp_infos(os.path.join(CMDIR, "tRNA5.c.cm"))["LI"].update_for_dbsize(Z)
    tol = e.lambda_ * 0.05 + math.log(1.05) + 1e-6
    for E, sc in rows:
        assert abs(math.log(float(e.evalue(sc)) / E)) <= tol, (E, sc, float(e.evalue(sc)))
    # a plausible slip (one strand only) fails
    e1 = cal.read_exp_infos(os.path.join(CMDIR, "tRNA5.c.cm"))["LI"].update_for_dbsize(Z / 2)
    assert any(abs(math.log(float(e1.evalue(sc)) / E)) > tol for E, sc in rows)


def test_score_for_evalue_inverts():
    e = cal.ExpInfo.from_fit(0.7, 3.0, 1.6e6, 250000, 0.004).update_for_dbsize(1e8)
    for E in (1e-10, 0.01, 1.0, 10.0):
        sc = e.score_for_evalue(E)
        assert abs(float(e.evalue(sc)) / E - 1) < 1e-9
    assert float(e.pvalue(e.mu_extrap - 5)) == 1.0


def test_fit_recovers_planted_exponential():
    rng = np.random.default_rng(3)
    lam, mu = 0.8, 2.0
    bulk = rng.normal(-5, 2, 300000)
    tail = mu + rng.exponential(1 / lam, 3000)
    x = np.concatenate([bulk, tail])
    tp = 2000 / len(x)
    m, l, n = cal.fit_exp_tail(x, tp)
    assert n == int(len(x) * tp)
    assert abs(l - lam) / lam < 0.08
    # the fit's mu is the smallest value kept
    assert m == np.sort(x)[len(x) - n]


def test_tail_mass_rule():
    assert abs(cal.tail_mass("LC", 1.6e6, 300199) - 0.003997) < 1e-6
    assert abs(cal.tail_mass("GC", 1.6e6, 77177) - 0.005183) < 1e-6
    with pytest.raises(ValueError):
        cal.tail_mass("LI", 1.6e6, 1000)


def test_genomic_sampler_composition():
    rng = np.random.default_rng(11)
    x = cal.sample_genomic(200, 10000, rng)
    T = np.asarray(cal._GHMM_TRANS)
    T = T / T.sum(1, keepdims=True)
    w, V = np.linalg.eig(T.T)
    pi = np.real(V[:, np.argmin(abs(w - 1))])
    pi /= pi.sum()
    E = np.asarray(cal._GHMM_EMIT)
    E = E / E.sum(1, keepdims=True)
    comp = pi @ E
    got = np.bincount(x.ravel(), minlength=4) / x.size
    # the chain is sticky (runs of ~thousands), so sampling error is large-ish
    assert np.abs(got - comp).max() < 0.03, (got, comp)


def test_cyk_filter_window_merge(monkeypatch):
    """pli_cyk_seq_filter's windows: stop-(W-1) .. start+(W-1), clipped,
    hits in position order, overlapping windows merged."""
    from infernal_amd import pipeline as pl

    class FakeCM:
        W = 10

    hits = [(50, 60, 0, 0, 5.0), (5, 12, 0, 0, 9.0), (64, 70, 0, 0, 4.0), (200, 210, 0, 0, 7.0)]
    seen = {}

    def fake_scan(cm, q, dsq, i0, j0, cutoff, do_null3=True, want_hits=True):
        seen["args"] = (q, i0, j0, cutoff)
        return {"hits": hits}

    monkeypatch.setattr(pl, "FastCYKScan", fake_scan)
    e = cal.ExpInfo.from_fit(0.7, 3.0, 1.6e6, 250000, 0.004).update_for_dbsize(1e6)
    wins = pl.cyk_filter_windows(FakeCM(), None, 205, e, F6=1e-4)
    # (5,12)->[3,14]; (50,60)->[51,59]; (64,70)->[61,73] does not touch [51,59]; (200,210)->[201,205]
    assert wins == [(3, 14), (51, 59), (61, 73), (201, 205)]
    q, i0, j0, cutoff = seen["args"]
    assert (q, i0, j0) == (1, 1, 205)
    assert abs(float(e.pvalue(cutoff)) - 1e-4) < 1e-12
    hits[:] = [(10, 20, 0, 0, 1.0), (15, 25, 0, 0, 1.0)]
    # merged: the window ends at the later hit's start + (W-1)
    assert pl.cyk_filter_windows(FakeCM(), None, 100, e) == [(11, 24)]


def test_revcomp():
    from infernal_amd import pipeline as pl
    assert pl.revcomp(np.array([0, 1, 2, 3, 3], dtype=np.uint8)).tolist() == [0, 0, 1, 2, 3]


def test_revcomp_degenerate_codes():
    """Every IUPAC letter's complement stands for the Watson-Crick
    complements of the residues the letter stands for (sets from the
    IUPAC-IUB table), the complement is an involution, and a target with N
    and other degenerate codes reverse-complements (the scanners reject codes
    above 15, so a wrong table would make the minus strand fail)."""
    from infernal_amd import pipeline as pl
    from test_host_model import IUPAC
    code = {ch: int(ia.digitize(ch)[1]) for ch in IUPAC}
    wc = {"A": "U", "C": "G", "G": "C", "U": "A"}
    by_set = {frozenset(IUPAC[ch]): code[ch] for ch in IUPAC}
    for ch in IUPAC:
        want = by_set[frozenset(wc[x] for x in IUPAC[ch])]
        assert pl.revcomp(np.array([code[ch]], dtype=np.uint8))[0] == want, ch
    allc = np.arange(18, dtype=np.uint8)
    assert np.array_equal(pl.revcomp(pl.revcomp(allc)), allc)
    seq = ia.digitize("ACGUNNRYKMSWBDHV")[1:-1]
    rc = pl.revcomp(seq)
    assert rc.max() <= 15 and len(rc) == len(seq)


def test_exp_lines_round_trip(tmp_path):
    src = os.path.join(CMDIR, "tRNA5.c.cm")
    infos = cal.read_exp_infos(src)
    text = cal.with_exp_lines(open(src).read(), infos)
    assert sum(1 for l in text.splitlines() if l.startswith("ECM")) == 4
    p = tmp_path / "x.cm"
    p.write_text(text)
    back = cal.read_exp_infos(str(p))
    for m in infos:
        a, b = infos[m], back[m]
        assert abs(a.lambda_ - b.lambda_) < 1e-5 and abs(a.mu_extrap - b.mu_extrap) < 1e-5
        assert a.nrandhits == b.nrandhits and abs(a.tailp - b.tailp) < 1e-6


def test_read_windows_cover_with_maxw_overlap():
    """esl_sqio_ReadWindow as cmsearch reads a sequence: windows of `window`
    new residues, each after the first starting maxW residues inside the
    previous one; every stretch of at most maxW residues lies inside a
    window."""
    from infernal_amd import pipeline as pl
    for L, maxW, window in ((1, 50, 100), (100, 50, 100), (101, 50, 100), (1000, 90, 100), (250000, 111, 100000)):
        ws = pl.read_windows(L, maxW, window)
        assert ws[0] == (1, min(L, window)) and ws[-1][1] == L
        for (a0, b0), (a1, b1) in zip(ws, ws[1:]):
            assert a1 == max(1, b0 - maxW + 1) and b1 == min(L, b0 + window)
        for s in range(1, L + 1, 7):
            e = min(L, s + maxW - 1)
            assert any(a <= s and e <= b for a, b in ws), (L, s, e)


def test_pipeline_passes_follow_terminus_rules():
    """The passes cm_Pipeline runs on a window (src/cm_pipeline.c:1375-1412,
    default truncation): standard always; 5' only with the 5' end, over the
    first maxW residues; 3' only with the 3' end, over the last maxW; 5' and
    3' with both ends and a window of at most maxW residues; none truncated
    with truncation off."""
    from infernal_amd import pipeline as pl
    P = ia.PLI_PASS_STD_ANY, ia.PLI_PASS_5P_ONLY_FORCE, ia.PLI_PASS_3P_ONLY_FORCE, ia.PLI_PASS_5P_AND_3P_FORCE
    assert pl.pipeline_passes(1000, False, False, 100) == [(P[0], 1, 1000)]
    assert pl.pipeline_passes(1000, True, False, 100) == [(P[0], 1, 1000), (P[1], 1, 100)]
    assert pl.pipeline_passes(1000, False, True, 100) == [(P[0], 1, 1000), (P[2], 901, 1000)]
    assert pl.pipeline_passes(1000, True, True, 100) == [(P[0], 1, 1000), (P[1], 1, 100), (P[2], 901, 1000)]
    assert pl.pipeline_passes(80, True, True, 100) == [(P[0], 1, 80), (P[1], 1, 80), (P[2], 1, 80), (P[3], 1, 80)]
    assert pl.pipeline_passes(80, True, True, 100, trunc=False) == [(P[0], 1, 80)]


def test_search_overlap_removal_best_first_per_strand():
    """Hits of one strand sharing a residue: the better one stays (equal
    scores: the lower start); the other strand is independent."""
    from infernal_amd import pipeline as pl
    hs = [{"start": 10, "stop": 50, "strand": "+", "score": 5.0}, {"start": 40, "stop": 90, "strand": "+", "score": 7.0},
          {"start": 91, "stop": 120, "strand": "+", "score": 1.0}, {"start": 60, "stop": 20, "strand": "-", "score": 3.0},
          {"start": 100, "stop": 55, "strand": "-", "score": 3.0}, {"start": 200, "stop": 150, "strand": "+", "score": 2.0}]
    kept = pl.remove_overlaps(hs)
    got = sorted((h["strand"], h["start"], h["stop"]) for h in kept)
    assert got == [("+", 40, 90), ("+", 91, 120), ("+", 200, 150), ("-", 60, 20)]


def test_search_overlap_removal_matches_quadratic_rule():
    """The pipeline's overlap removal on random hits with ties, both strands,
    against the quadratic scan it restates (a hit is dropped if it shares a
    residue with any hit kept before it, in score-then-start order)."""
    from infernal_amd import pipeline as pl
    rng = np.random.default_rng(77)
    hs = []
    for _ in range(3000):
        a = int(rng.integers(1, 20000))
        b = a + int(rng.integers(0, 300))
        strand = "+" if rng.random() < 0.5 else "-"
        start, stop = (a, b) if strand == "+" else (b, a)
        hs.append({"start": start, "stop": stop, "strand": strand, "score": float(rng.integers(0, 40)) / 4.0})
    got = pl.remove_overlaps(hs)
    want = []
    for strand in ("+", "-"):
        cand = sorted((h for h in hs if h["strand"] == strand), key=lambda h: (-h["score"], h["start"]))
        kept = []
        for h in cand:
            lo, hi = min(h["start"], h["stop"]), max(h["start"], h["stop"])
            if all(hi < min(k["start"], k["stop"]) or lo > max(k["start"], k["stop"]) for k in kept):
                kept.append(h)
        want += kept
    assert [id(h) for h in got] == [id(h) for h in want]
    assert len(got) > 50
