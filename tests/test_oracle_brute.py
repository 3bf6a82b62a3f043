"""Pin the oracle's standard scanners against brute-force parse enumeration.

For tiny random models (chains, pairs, bifurcations; local and global), every
root cell bestsc[d] at every j that the oracle reports must equal the best
(CYK) or summed (Inside) score over all parse trees of x[j-d+1..j], as listed
by explicit enumeration.  Tolerances follow src/itest_brute.c:280-318 (the
enumeration adds the same float scores in a different order): 1e-3 bits for
CYK, 0.015 bits for Inside, 0.02 bits for the integer Inside (its log-sums
are rounded to 1/1000 bit at every step).
"""
import numpy as np
import pytest

import infernal_amd as ia
import oracle_bind as ob
from tinycm import make_cm_text, Tables, root_score

LAYOUTS = [
    ["ROOT", "MATL", "MATL", "END"],
    ["ROOT", "MATP", "MATL", "END"],
    ["ROOT", "MATR", "MATP", "END"],
    ["ROOT", "BIF", "BEGL", "MATL", "END", "BEGR", "MATL", "END"],
    ["ROOT", "MATP", "BIF", "BEGL", "MATR", "END", "BEGR", "MATL", "END"],
]


def _check(layout, seed, local, L):
    text = make_cm_text(layout, seed, W=8)
    cm = ia.CM.from_text(text)
    cm.configure(local=local)
    tb = Tables(cm)
    rng = np.random.default_rng(seed + 1000)
    codes = rng.integers(0, 4, L).astype(np.uint8)
    dsq = ia.dsq_from_codes(codes)
    x = [0] + [int(c) for c in codes]
    for algo, inside, tol in ((ob.ORACLE_CYK, False, 1e-3), (ob.ORACLE_FINSIDE, True, 0.015), (ob.ORACLE_IINSIDE, True, 0.02)):
        o = ob.scan(cm, algo, 0, dsq, 1, L, 1e30, want_rows=True)
        W = o["W"]
        for j in range(1, L + 1):
            for d in range(1, min(j, W) + 1):
                got = float(o["rows_sc"][j - 1][d])
                want = root_score(tb, x, j - d + 1, j, inside)
                if want is None or want < -1e29:
                    assert got < -9e35, (layout, algo, j, d, got)
                else:
                    assert abs(got - want) <= tol, (layout, local, algo, j, d, got, want)


@pytest.mark.parametrize("layout", LAYOUTS, ids=["chain", "pair", "matr", "bif", "pairbif"])
@pytest.mark.parametrize("local", [True, False], ids=["local", "glocal"])
def test_oracle_vs_brute(layout, local):
    for seed in (1, 2, 3):
        _check(layout, seed, local, L=5)
