"""Synthetic inputs: homolog emission and the Markov background.

Emission (infernal_amd.synth.emit, following EmitParsetree,
src/cm_parsetree.c:1890) is checked against what the model fixes in closed
form: the expected emitted length E[v] = (residues state v emits) +
sum_k t[v][k] E[cfirst[v]+k], with E[B] = E[left] + E[right] and E[E] = 0,
must match the sample mean within four standard errors; every emitted
sequence of a model without insert-state detours has at least the
consensus-pair count of residues.  The background must have the requested
GC content and a real second-order dependence.
"""
import os

import numpy as np

import infernal_amd as ia
from infernal_amd import synth

CMDIR = os.path.join(os.path.dirname(__file__), "golden", "cm")
DELTA = {ia.MP_st: 2, ia.ML_st: 1, ia.MR_st: 1, ia.IL_st: 1, ia.IR_st: 1}


def expected_length(cm):
    c = cm.c
    E = [0.0] * c.M
    for v in range(c.M - 1, -1, -1):
        s = ord(c.sttype[v])
        if s == ia.E_st:
            continue
        if s == ia.B_st:
            E[v] = E[c.cfirst[v]] + E[c.cnum[v]]
            continue
        n = c.cnum[v]
        t = np.array([c.t[v][k] for k in range(n)], dtype=np.float64)
        t /= t.sum()
        # the self-loop of an insert state: E = (delta + sum_{k>0} t_k E_k) / (1 - t_0)
        if s in (ia.IL_st, ia.IR_st) and c.cfirst[v] == v:
            E[v] = (DELTA[s] + sum(t[k] * E[v + k] for k in range(1, n))) / (1.0 - t[0])
        else:
            E[v] = DELTA.get(s, 0) + sum(t[k] * E[c.cfirst[v] + k] for k in range(n))
    return E[0]


def test_emitted_length_matches_model():
    for name, n in (("tRNA5.c.cm", 3000), ("synth_cloverleaf76.cm", 3000)):
        cm = ia.CM.read(os.path.join(CMDIR, name))
        rng = np.random.default_rng(5)
        lens = np.array([len(synth.emit(cm, rng)) for _ in range(n)], dtype=np.float64)
        mu = expected_length(cm)
        se = lens.std(ddof=1) / np.sqrt(n)
        assert abs(lens.mean() - mu) < 4 * se + 1e-9, (name, lens.mean(), mu, se)


def test_emitted_homologs_are_fragments():
    cm = ia.CM.read(os.path.join(CMDIR, "synth_cloverleaf76.cm"))
    frags = synth.homologs(cm, 50, seed=3, frac=(0.5, 1.0))
    assert len(frags) == 50
    assert all(f.dtype == np.uint8 and f.max() <= 3 for f in frags)
    assert min(len(f) for f in frags) >= 30


def test_markov_background_composition():
    x = synth.markov2_background(400_000, gc=0.40, seed=9)
    assert x.dtype == np.uint8 and x.max() <= 3
    gc = np.isin(x, (1, 2)).mean()
    assert abs(gc - 0.40) < 0.01
    # second order: P(x_t | x_{t-2}, x_{t-1}) differs across contexts
    ctx = x[:-2].astype(np.int64) * 4 + x[1:-1]
    nxt = x[2:]
    freq = np.array([np.bincount(nxt[ctx == c], minlength=4) / max(1, (ctx == c).sum()) for c in range(16)])
    assert freq.std(axis=0).max() > 0.02


def test_plant_places_fragments():
    bg = np.zeros(10_000, dtype=np.uint8)
    frags = [np.full(100, 1, dtype=np.uint8), np.full(50, 2, dtype=np.uint8)]
    codes, starts = synth.plant(bg, frags, seed=4)
    for s, f in zip(starts, frags):
        assert np.array_equal(codes[s - 1:s - 1 + len(f)], f)
    assert (codes != 0).sum() == 150
