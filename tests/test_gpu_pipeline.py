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
        hits, _ = pl.search(CMF, codes, Z=5874406, E=10.0, cm=cm)
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
