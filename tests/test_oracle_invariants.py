"""Invariants of the oracle's standard scanners on shipped models (CPU).

* Inside >= CYK cell by cell (FLogsum(a,b) >= max(a,b); table entries >= 0).
* Band nesting: QDB1 bands lie inside QDB2 bands, which lie inside no bands,
  so root cells that exist in the tighter set can only grow.
* Chunk invariance: a scan started W-1 residues before j reproduces the root
  row at j exactly (a cell only reads residues j-d+1..j).  The GPU's chunked
  launch relies on this.
"""
import numpy as np
import pytest

import oracle_bind as ob
from helpers import load_cm, random_dsq


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Vault.c.cm", "Cobalamin.c.cm"])
def test_inside_ge_cyk(name):
    cm = load_cm(name)
    L = 600
    dsq = random_dsq(L, seed=2)
    c = ob.scan(cm, ob.ORACLE_CYK, 1, dsq, 1, L, 1e30, want_rows=True)["rows_sc"]
    f = ob.scan(cm, ob.ORACLE_FINSIDE, 1, dsq, 1, L, 1e30, want_rows=True)["rows_sc"]
    i = ob.scan(cm, ob.ORACLE_IINSIDE, 1, dsq, 1, L, 1e30, want_rows=True)["rows_sc"]
    live = c > -9e35
    assert np.all(f[live] >= c[live])
    assert np.all(i[live] >= c[live] - 0.01 * 40)   # integer rounding, bounded by the number of terms
    assert np.all(np.abs(f[live] - i[live]) < 0.5)


@pytest.mark.parametrize("name", ["tRNA.c.cm", "5S_rRNA.c.cm"])
def test_band_nesting(name):
    cm = load_cm(name)
    L = 500
    dsq = random_dsq(L, seed=4)
    r = {q: ob.scan(cm, ob.ORACLE_CYK, q, dsq, 1, L, 1e30, want_rows=True)["rows_sc"] for q in (0, 1, 2)}
    for tight, loose in ((1, 2), (2, 0)):
        live = r[tight] > -9e35
        assert np.all(r[loose][live] >= r[tight][live])


@pytest.mark.parametrize("algo", [ob.ORACLE_CYK, ob.ORACLE_FINSIDE, ob.ORACLE_IINSIDE])
def test_chunk_invariance(algo):
    cm = load_cm("tRNA.c.cm")
    W = cm.W
    L = 1500
    dsq = random_dsq(L, seed=6)
    full = ob.scan(cm, algo, 1, dsq, 1, L, 1e30, want_rows=True)
    a = 700
    s = a - W + 1
    part = ob.scan(cm, algo, 1, dsq, s, L, 1e30, want_rows=True)
    assert part["W"] == full["W"]
    fr = full["rows_sc"][a - 1:]
    pr = part["rows_sc"][a - s:]
    assert np.array_equal(fr.view(np.uint32), pr.view(np.uint32))
    assert np.array_equal(full["rows_r"][a - 1:], part["rows_r"][a - s:])
