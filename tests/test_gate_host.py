"""Host logic of the learned gate (no GPU): window selection, recall,
labels and bf16 rounding."""
import numpy as np

from infernal_amd import gate as g


def _windows_brute(logits, tau, W):
    n = len(logits)
    cover = np.zeros(n + 2, dtype=bool)
    for t in np.nonzero(logits > tau)[0] + 1:
        cover[max(1, t - (W - 1)):min(n, t + (W - 1)) + 1] = True
    out, t = [], 1
    while t <= n:
        if cover[t]:
            s = t
            while t + 1 <= n and cover[t + 1]:
                t += 1
            out.append((s, t))
        t += 1
    return out


def test_select_windows_matches_brute_force():
    rng = np.random.default_rng(1)
    for trial in range(50):
        n = int(rng.integers(1, 400))
        logits = rng.standard_normal(n).astype(np.float32)
        tau = float(rng.uniform(0.5, 3.0))
        W = int(rng.integers(1, 40))
        assert g.select_windows(logits, tau, W) == _windows_brute(logits, tau, W), (n, tau, W)


def test_recall_and_labels():
    plain = [(10, 50, 30.0), (200, 150, 40.0), (400, 480, 25.0)]
    gated = [(12, 52, 29.9), (150, 200, 40.0)]
    assert abs(g.recall(plain, gated) - 2 / 3) < 1e-12
    assert g.recall(plain, gated + [(470, 500, 5.0)]) == 2 / 3          # overlap below half
    assert g.recall(plain, gated + [(420, 480, 5.0)]) == 1.0
    lab = g.labels_from_hits(500, plain)
    assert lab.sum() == 41 + 51 + 81 and lab[9] == 1 and lab[8] == 0 and lab[149] == 1


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -2.5, 3.14159, 1e-20], dtype=np.float32)
    h = g.to_bf16_bits(x)
    back = g.from_bf16_bits(h)
    # ties to even: 1 + 2^-8 -> 1.0, 1 + 3*2^-8 -> 1 + 2^-6
    assert back[1] == 1.0 and back[3] == 1.015625
    assert np.all(np.abs(back - x) <= np.abs(x) * 2.0 ** -8)


def test_smooth_is_a_centred_mean():
    rng = np.random.default_rng(2)
    x = rng.standard_normal(300).astype(np.float32)
    s = g.smooth(x, 16)
    for t in (0, 5, 100, 299):
        lo, hi = max(0, t - 8), min(300, t + 8)
        assert abs(s[t] - x[lo:hi].mean()) < 1e-5


def _windows_brute_m(logits, tau, m):
    """_windows_brute with an explicit widening m (W - 1 = m)."""
    return _windows_brute(logits, tau, m + 1)


def test_select_windows_widening_and_cover():
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(1, 400))
        logits = rng.standard_normal(n).astype(np.float32)
        tau = float(rng.uniform(0.5, 3.0))
        W = int(rng.integers(2, 60))
        m = int(rng.integers(0, W))
        wins = g.select_windows(logits, tau, W, m)
        assert wins == _windows_brute_m(logits, tau, m), (n, tau, W, m)
        assert g.select_windows(logits, tau, W, 10 * W) == g.select_windows(logits, tau, W)   # capped at W - 1
        hits = []
        for _ in range(20):
            a = int(rng.integers(1, n + 1)); b = int(rng.integers(a, min(n, a + 30) + 1))
            hits.append((a, b) if rng.random() < 0.5 else (b, a))
        got = g.windows_cover(wins, hits)
        for h, ok in zip(hits, got):
            lo, hi = min(h), max(h)
            assert ok == any(s <= lo and hi <= e for s, e in wins)


def test_fit_widen_takes_the_smallest_grid_widening_that_keeps_the_hits():
    n, W = 8000, 1500
    logits = np.zeros(n, dtype=np.float32)
    hits = []
    # hits whose flagged core misses up to 40 residues at either end
    for k, start in enumerate(range(100, 4800, 400)):
        a, b = start, start + 200
        logits[a - 1 + 40 - (k % 3) * 10:b - 40 + (k % 2) * 10] = 5.0
        hits.append((a, b))
    assert g.fit_widen(logits, 1.0, hits, W) == 40           # cores end up to 40 short of their hits
    assert g.windows_cover(g.select_windows(logits, 1.0, W, 40), hits).all()
    assert not g.windows_cover(g.select_windows(logits, 1.0, W, 39), hits).all()
    assert g.fit_widen(logits, 1.0, [], W) == W - 1
    # a hit more than W - 1 from every flagged residue (no window holds it at
    # any widening) does not force W - 1; one that only a wide window
    # reaches (the last core ends at 4670) needs 5590 - 4670
    assert g.fit_widen(logits, 1.0, hits + [(7800, 7890)], W) == 40
    assert g.fit_widen(logits, 1.0, hits + [(5500, 5590)], W) == 920
    # the slack: one hit in 13 above 0.002 is not given up ...
    assert g.fit_widen(logits, 1.0, hits + [(5500, 5590)], W, slack=0.05) == 920
    # ... one in 13 below 0.1 is
    assert g.fit_widen(logits, 1.0, hits + [(5500, 5590)], W, slack=0.1) == 40
