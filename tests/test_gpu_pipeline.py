"""The CM-only search pipeline (cmsearch --nohmm) on the GPU scanners vs the
reference's tutorial run: tRNA5.c.cm against mrum-genome.fa
(tutorial/tRNA5-mrum.out, Z = 5,874,406 residues).  The 26 genome stretches
around the tutorial's 56 hits are searched on both strands with the E-values
of the full search space.  The tutorial run used HMM-banded Inside; this
pipeline uses QDB-banded Inside, which cannot return hits longer than the
model's W (89).  Bar: every tutorial hit of length <= W is found on its
strand with the same start and stop, its score within 0.15 bit (the
output prints 0.1-bit scores) and its E-value within 15%."""
import os

import numpy as np
import pytest

import infernal_amd as ia
from infernal_amd import pipeline as pl

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CMF = os.path.join(GOLD, "cm", "tRNA5.c.cm")


def _regions():
    recs = []
    with open(os.path.join(GOLD, "mrum_hit_regions.fa")) as fh:
        for line in fh:
            if line.startswith(">"):
                a, b = map(int, line[6:].split("_"))
                recs.append([a, b, []])
            else:
                recs[-1][2].append(line.strip())
    out = []
    for a, b, parts in recs:
        s = "".join(parts).upper().replace("T", "U")
        out.append((a, b, np.array(["ACGU".index(c) if c in "ACGU" else 0 for c in s], dtype=np.uint8)))
    return out


def _reference():
    rows = []
    with open(os.path.join(GOLD, "tRNA5_mrum_hits.tsv")) as fh:
        for line in fh:
            if not line.startswith("#"):
                f = line.rstrip("\n").split("\t")
                rows.append((float(f[1]), float(f[2]), int(f[3]), int(f[4]), f[5]))
    return rows


def test_tutorial_search_nohmm():
    cm = ia.CM.read(CMF)
    cm.configure(local=True)
    found = []
    for a, b, codes in _regions():
        print("case pipeline region", a, b, flush=True)
        hits, _ = pl.search(CMF, codes, Z=5874406, E=10.0, cm=cm, trunc=False)   # stretches, not whole sequences
        for h in hits:
            found.append((h["start"] + a - 1, h["stop"] + a - 1, h["strand"], h["score"], h["evalue"]))
    W = cm.W
    n_checked = 0
    for E, sc, s, e, strand in _reference():
        if abs(e - s) + 1 > W:
            continue
        m = [f for f in found if f[2] == strand and f[0] == s and f[1] == e]
        assert m, ("missed", E, sc, s, e, strand)
        assert abs(m[0][3] - sc) <= 0.15, (m[0], sc)
        assert abs(m[0][4] / E - 1) <= 0.15, (m[0], E)
        n_checked += 1
    assert n_checked >= 55


def _planted(L, positions, seed, clip5=0):
    """Background of L residues with tRNA5 homologs starting at the given
    positions (the first one's first clip5 residues cut away)."""
    from infernal_amd import synth
    rng = np.random.default_rng(seed)
    seq = synth.markov2_background(L, gc=0.45, seed=seed)
    src = ia.CM.read(CMF)
    for k, p in enumerate(positions):
        h = synth.emit(src, rng)
        if k == 0 and clip5:
            h = h[clip5:]
        seq[p - 1:p - 1 + len(h)] = h
    return seq


def test_windowed_search_equals_one_window():
    """cmsearch's 100-kb windows with maxW overlap: a 230-kb sequence with
    homologs inside windows and across both window boundaries returns the
    same E <= 0.01 hits (coordinates, strand, score bits) as one window over
    the whole sequence, truncated passes at both ends included."""
    cm = ia.CM.read(CMF)
    cm.configure(local=True, trunc=True, scanmx=True, trscanmx=True)
    L = 230000
    codes = _planted(L, [5000, 99970, 150000, 199950, 225000], seed=31)
    print("case windowed search", L, flush=True)
    a, sa = pl.search(CMF, codes, E=0.01, cm=cm)
    b, sb = pl.search(CMF, codes, E=0.01, cm=cm, window=10 ** 9)
    assert sa["+"]["windows"] == 3 and sb["+"]["windows"] == 1
    key = lambda h: (h["strand"], h["start"], h["stop"])
    ka = sorted(((h["strand"], h["start"], h["stop"], np.float32(h["score"]).view(np.uint32)) for h in a))
    kb = sorted(((h["strand"], h["start"], h["stop"], np.float32(h["score"]).view(np.uint32)) for h in b))
    assert ka == kb, (ka, kb)
    assert len(ka) >= 5


def test_truncated_pass_finds_clipped_five_prime_end():
    """A sequence that begins inside a homolog (its first 25 residues cut
    away): the 5'-truncation pass reports it starting at residue 1, above the
    best score of the pipeline without truncated passes there."""
    cm = ia.CM.read(CMF)
    cm.configure(local=True, trunc=True, scanmx=True, trscanmx=True)
    codes = _planted(6000, [1], seed=33, clip5=25)
    print("case truncated 5' end", flush=True)
    with_t, _ = pl.search(CMF, codes, E=10.0, cm=cm, both_strands=False)
    without, _ = pl.search(CMF, codes, E=10.0, cm=cm, both_strands=False, trunc=False)
    t5 = [h for h in with_t if h["pass"] == ia.PLI_PASS_5P_ONLY_FORCE]
    assert t5 and t5[0]["start"] == 1, with_t[:3]
    near = [h["score"] for h in without if min(h["start"], h["stop"]) <= 80]
    assert not near or t5[0]["score"] > max(near), (t5[0], near)
