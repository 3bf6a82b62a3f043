"""Pin the oracle's truncated scanners (RefTrCYKScan, RefITrInsideScan).

1. The truncated-parse enumerator (tests/tinycm_trunc.py) lists exactly the
   Joint, Left- and Right-marginal paths that the reference's own brute-force
   test enumerates by hand for its ROOT-MATL-MATL-END model and sequences of
   one and two residues (src/itest_brute.c score_brute_matl_cm; fixture
   tests/golden/itest_brute_matl_paths.json, extracted by
   tests/golden/make_itest_brute_paths.py).
2. For tiny random models (chains, pairs, MATR, bifurcations; local and
   glocal; the three truncation passes), every per-mode root cell the oracle
   computes -- max over truncated begins y of X_mode[y][d] + penalty[y] --
   equals the best (TrCYK, 1e-3 bits) or summed (TrInside, 0.02 bits) score
   over the enumerated truncated parses of x[j-d+1..j].
3. The root row bestsc/bestmode/bestr follows from those per-mode cells by
   the reference's record rule (src/cm_dpsearch_trunc.c:538-605): the root
   cells X_mode[cur][0][d] are never reset, so a mode's best begin at j only
   counts if it beats that mode's best at every earlier j of the same parity.
"""
import json
import os
from collections import Counter

import numpy as np
import pytest

import infernal_amd as ia
import oracle_bind as ob
from tinycm import make_cm_text
from tinycm_trunc import TrTables, mode_maxima, symbolic_paths, TRMODE_J, TRMODE_L, TRMODE_R, TRMODE_T

HERE = os.path.dirname(os.path.abspath(__file__))
MI = {"J": TRMODE_J, "L": TRMODE_L, "R": TRMODE_R, "T": TRMODE_T}
ORDER = ("J", "L", "R", "T")
# pass index -> (modes filled, penalty index)  (src/cm_dpsearch_trunc.c:3566-3587, src/cm_trunc.c:464-473)
PASSES = {ia.PLI_PASS_5P_AND_3P_FORCE: ("JLRT", 0), ia.PLI_PASS_5P_ONLY_FORCE: ("JR", 1),
          ia.PLI_PASS_3P_ONLY_FORCE: ("JL", 2)}
LAYOUTS = [
    ["ROOT", "MATL", "MATL", "END"],
    ["ROOT", "MATP", "MATL", "END"],
    ["ROOT", "MATR", "MATP", "END"],
    ["ROOT", "BIF", "BEGL", "MATL", "END", "BEGR", "MATL", "END"],
    ["ROOT", "MATP", "BIF", "BEGL", "MATR", "END", "BEGR", "MATL", "END"],
]
IMPOSSIBLE_BELOW = -1e5


def test_enumerator_lists_the_reference_hand_paths():
    ref = json.load(open(os.path.join(HERE, "golden", "itest_brute_matl_paths.json")))
    cm = ia.CM.from_text(make_cm_text(ref["model"], 1, W=8))
    cm.configure(local=True, trunc=True)
    tb = TrTables(cm, 0)
    # the reference model: local end only from MATL_ML(3), state 8 detached
    assert [v for v in range(tb.M) if tb.endsc[v] > -1e30] == [3]
    assert [v for v in range(tb.M) if tb.detached(v)] == [8]
    for L in (1, 2):
        got = symbolic_paths(tb, L)
        for m in "JLR":
            want = Counter(tuple(sorted(p)) for p in ref["paths"][m][str(L)])
            assert got[m] == want, (m, L, list((want - got[m]).elements()), list((got[m] - want).elements()))


def _record_rule(msc, modes, j0_abs):
    """bestsc/bestmode per (j, d) from per-mode cells by the reference's
    root update, with its running per-parity maxima."""
    L, W1, _ = msc.shape
    run = {p: {m: np.full(W1, -np.inf, dtype=np.float64) for m in modes} for p in (0, 1)}
    best = np.full((L, W1), -np.inf)
    bmode = np.full((L, W1), -1)
    for r in range(L):
        par = (j0_abs + r) % 2
        for d in range(1, W1):
            for m in ORDER:
                if m not in modes:
                    continue
                v = float(msc[r, d, MI[m]])
                if v <= IMPOSSIBLE_BELOW:
                    continue
                if v > run[par][m][d]:
                    run[par][m][d] = v
                    if v > best[r, d]:
                        best[r, d], bmode[r, d] = v, MI[m]
    return best, bmode


def _check(layout, seed, local, pass_idx, L=5):
    modes, pty = PASSES[pass_idx]
    cm = ia.CM.from_text(make_cm_text(layout, seed, W=8))
    cm.configure(local=local, trunc=True)
    tb = TrTables(cm, pty)
    rng = np.random.default_rng(seed + 1000)
    codes = rng.integers(0, 4, L).astype(np.uint8)
    dsq = ia.dsq_from_codes(codes)
    x = [0] + [int(c) for c in codes]
    for algo, inside, tol in ((ob.ORACLE_TRCYK, False, 1e-3), (ob.ORACLE_TRIINSIDE, True, 0.02)):
        o = ob.scan(cm, algo, 0, dsq, 1, L, 1e30, want_rows=True, pass_idx=pass_idx)
        W = o["W"]
        msc = o["rows_msc"]
        for j in range(1, L + 1):
            for d in range(1, min(j, W) + 1):
                mm = mode_maxima(tb, x, j - d + 1, j, modes, inside)
                for m in modes:
                    got = float(msc[j - 1][d][MI[m]])
                    want = mm[m][0]
                    ctx = (layout, local, pass_idx, algo, j, d, m, got, want)
                    if want is None:
                        assert got <= IMPOSSIBLE_BELOW, ctx
                    else:
                        assert abs(got - want) <= tol, ctx
                # the root's state: a begin whose own enumerated score attains its mode's best
                r, bm = int(o["rows_r"][j - 1][d]), int(o["rows_mode"][j - 1][d])
                if bm in MI.values():
                    m = next(k for k, v in MI.items() if v == bm)
                    assert r in mm[m][2] and abs(mm[m][2][r] - mm[m][0]) <= 2 * tol, (layout, algo, j, d, m, r)
        best, bmode = _record_rule(msc, modes, 1)
        for j in range(1, L + 1):
            for d in range(1, min(j, W) + 1):
                got = float(o["rows_sc"][j - 1][d])
                if best[j - 1, d] == -np.inf:
                    assert got <= IMPOSSIBLE_BELOW, (layout, algo, j, d, got)
                    assert int(o["rows_mode"][j - 1][d]) == 4   # TRMODE_UNKNOWN
                else:
                    assert got == np.float32(best[j - 1, d]), (layout, algo, j, d, got, best[j - 1, d])
                    assert int(o["rows_mode"][j - 1][d]) == bmode[j - 1, d]


@pytest.mark.parametrize("layout", LAYOUTS, ids=["chain", "pair", "matr", "bif", "pairbif"])
@pytest.mark.parametrize("local", [True, False], ids=["local", "glocal"])
@pytest.mark.parametrize("pass_idx", list(PASSES), ids=["5and3", "5only", "3only"])
def test_trunc_oracle_vs_brute(layout, local, pass_idx):
    for seed in (1, 2, 3):
        _check(layout, seed, local, pass_idx)


def test_record_rule_drops_later_lower_begins():
    """The root cells persist across rows of one parity: a cell whose best
    begin is lower than at j-2 reports nothing in that mode."""
    msc = np.full((4, 2, 4), -1e30, dtype=np.float32)
    msc[0, 1, TRMODE_J] = -1.0    # j = 1
    msc[2, 1, TRMODE_J] = -2.0    # j = 3: lower than at j = 1
    msc[2, 1, TRMODE_L] = -3.0    #        but L is new
    msc[1, 1, TRMODE_J] = -5.0    # j = 2: other parity
    best, bmode = _record_rule(msc, "JL", 1)
    assert best[0, 1] == -1.0 and bmode[0, 1] == TRMODE_J
    assert best[2, 1] == -3.0 and bmode[2, 1] == TRMODE_L
    assert best[1, 1] == -5.0
    assert best[3, 1] == -np.inf
