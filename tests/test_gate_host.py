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
