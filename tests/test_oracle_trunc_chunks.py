"""A truncated scan can be split into chunks of one window.

The truncated scanners keep one running maximum per (mode, row parity, d)
at the root that is never reset (src/cm_dpsearch_trunc.c:538-605), so a
window is scanned by one wavefront today.  This pins the two facts a chunked
scan would rest on, with the oracle on the CPU:

1. chunk invariance of the DP: the per-mode root maxima of rows a..b
   computed from a - W + 1 (W - 1 rows of warm-up) equal those of the whole
   window, bit for bit;
2. the root's record rule composes over chunks: applying it to a chunk's
   rows with the chunk's own running maxima, each seeded with the maxima of
   all earlier chunks (per mode, parity and d), gives the whole window's
   root row: bestsc bit for bit, and the mode wherever no other mode ties
   its score (the root visits begins state by state and modes within a
   state, so a tie between modes goes to the lower begin state; per-mode
   maxima alone do not carry it -- a chunked kernel keeps (score, state)
   keys).
"""
import numpy as np
import pytest

import infernal_amd as ia
import oracle_bind as ob
from helpers import load_cm, random_dsq

MODES = (3, 2, 1, 0)   # TRMODE_J, L, R, T: the order the root visits them


def _rule(msc, j_first, carry):
    """The root's record rule over rows j_first.. of msc [rows][d][mode];
    carry[parity][mode][d] holds the running maxima from earlier rows and is
    updated in place.  Returns bestsc [rows][d] and bestmode."""
    rows, W1, _ = msc.shape
    best = np.full((rows, W1), -np.inf)
    mode = np.full((rows, W1), 4)
    for r in range(rows):
        p = (j_first + r) % 2
        for d in range(1, W1):
            for m in MODES:
                v = float(msc[r, d, m])
                if v <= -1e5:
                    continue
                if v > carry[p][m][d]:
                    carry[p][m][d] = v
                    if v > best[r, d]:
                        best[r, d], mode[r, d] = v, m
    return best, mode


@pytest.mark.parametrize("algo", [ob.ORACLE_TRCYK, ob.ORACLE_TRIINSIDE], ids=["trcyk", "triinside"])
def test_chunked_truncated_scan_reproduces_the_window(algo):
    cm = load_cm("tRNA.c.cm", local=True, trunc=True)
    L, C = 1400, 450
    dsq = random_dsq(L, seed=77)
    qdb = 1
    p = ia.PLI_PASS_5P_AND_3P_FORCE
    full = ob.scan(cm, algo, qdb, dsq, 1, L, 1e30, want_rows=True, pass_idx=p)
    W = full["W"]
    assert C >= 2 * W
    carry = [[np.full(W + 1, -np.inf) for _ in range(4)] for _ in range(2)]
    ties = 0
    for a in range(1, L + 1, C):
        b = min(L, a + C - 1)
        s0 = max(1, a - W + 1)
        ch = ob.scan(cm, algo, qdb, dsq, s0, b, 1e30, want_rows=True, pass_idx=p)
        assert ch["W"] == W
        got = ch["rows_msc"][a - s0:]
        want = full["rows_msc"][a - 1:b]
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (a, b)
        # the record rule, seeded with the earlier chunks' maxima
        best, mode = _rule(got, a, carry)
        fs, fm = full["rows_sc"][a - 1:b], full["rows_mode"][a - 1:b]
        for r in range(b - a + 1):
            for d in range(1, min(a + r, W) + 1):
                if best[r, d] == -np.inf:
                    assert fs[r, d] <= -1e5 and fm[r, d] == 4, (a + r, d)
                else:
                    assert np.float32(best[r, d]) == fs[r, d], (a + r, d, best[r, d], fs[r, d])
                    if mode[r, d] != fm[r, d]:   # only a tie between modes may differ
                        assert np.float32(got[r, d, fm[r, d]]) == fs[r, d], (a + r, d, mode[r, d], fm[r, d])
                        ties += 1
    print("mode ties", ties)
