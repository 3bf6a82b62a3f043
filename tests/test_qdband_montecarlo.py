"""The band engine's length densities against sampled parse lengths, the
method of the reference's bandcyk-montecarlo-test (src/bandcyk-montecarlo-
test.c: sample parse trees, compare the length histogram with
BandCalculationEngine's gamma by chi-square).

For a state v, gamma_v(n) (infx_BandCalculationEngine's ret_gamma) is the
probability that the sub-model rooted at v emits n residues.  Here lengths
are sampled from the model's transition probabilities directly -- a walk
from v that counts emissions (2 for a pair state, 1 for a singlet, 0 for
D/S/B/E) and descends both children of a bifurcation -- independent of the
engine's convolution recursions.  Internal states only (bifurcations, their
BEGL/BEGR starts, pair states below them): the root's density also depends
on how the engine treats local begins.  Seeded, so the test is deterministic;
the bar is a chi-square p-value of 1e-4 over bins of expected count >= 5.
"""
import os

import numpy as np
import pytest
from scipy import stats

import infernal_amd as ia
from helpers import CMDIR
from test_qdband import engine_bands

S_st, E_st, B_st = 6, 7, 8          # state types (src/infernal.h)
MP_st = 1
EMITS = {0: 0, 1: 2, 2: 1, 3: 1, 4: 1, 5: 1, S_st: 0, E_st: 0, B_st: 0}   # D MP ML MR IL IR S E B


def sample_lengths(cm, v0, n, rng):
    c = cm.c
    st = cm.sttype()
    M = c.M
    cdf = {}
    for v in range(M):
        if st[v] in (E_st, B_st):
            continue
        p = np.array([c.t[v][k] for k in range(c.cnum[v])], dtype=np.float64)
        cdf[v] = np.cumsum(p / p.sum())
    out = np.zeros(n, dtype=np.int64)
    for r in range(n):
        total, stack = 0, [v0]
        while stack:
            v = stack.pop()
            s = st[v]
            if s == E_st:
                continue
            if s == B_st:
                stack.append(c.cfirst[v])
                stack.append(c.cnum[v])
                continue
            total += EMITS[s]
            stack.append(c.cfirst[v] + int(np.searchsorted(cdf[v], rng.random(), side="right")))
        out[r] = total
    return out


def chi2_pvalue(lengths, gamma):
    n = len(lengths)
    obs = np.bincount(lengths, minlength=len(gamma)).astype(np.float64)[:len(gamma)]
    assert (lengths < len(gamma)).all()
    exp = gamma / gamma.sum() * n
    # pool adjacent bins (in length order) until each expected count is >= 5
    o_b, e_b, o_acc, e_acc = [], [], 0.0, 0.0
    for o, e in zip(obs, exp):
        o_acc += o
        e_acc += e
        if e_acc >= 5.0:
            o_b.append(o_acc); e_b.append(e_acc); o_acc = e_acc = 0.0
    if e_acc > 0:
        o_b[-1] += o_acc; e_b[-1] += e_acc
    o_b, e_b = np.array(o_b), np.array(e_b)
    chi2 = ((o_b - e_b) ** 2 / e_b).sum()
    return stats.chi2.sf(chi2, len(o_b) - 1), len(o_b)


def picked_states(cm):
    st = cm.sttype()
    c = cm.c
    bifs = [v for v in range(c.M) if st[v] == B_st]
    states = []
    for b in bifs:
        states += [b, c.cfirst[b], c.cnum[b]]
    mps = [v for v in range(c.M) if st[v] == MP_st]
    states += mps[len(mps) // 2:len(mps) // 2 + 2]
    return sorted(set(states))


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Cobalamin.c.cm"])
def test_band_densities_match_sampled_lengths(name):
    cm = ia.CM.read(os.path.join(CMDIR, name))          # global model, the file's probabilities
    _, W, Z, gamma = engine_bands(cm, cm.c.beta_W)
    rng = np.random.default_rng(2024)
    checked = 0
    for v in picked_states(cm):
        lengths = sample_lengths(cm, v, 4000, rng)
        p, nb = chi2_pvalue(lengths, gamma[v])
        assert nb >= 5, (name, v, nb)
        assert p > 1e-4, (name, v, p, nb)
        checked += 1
    assert checked >= 5


def test_chi2_rejects_a_shifted_density():
    """The same comparison fails when the density is off by one residue."""
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm"))
    _, W, Z, gamma = engine_bands(cm, cm.c.beta_W)
    st = cm.sttype()
    b = [v for v in range(cm.c.M) if st[v] == B_st][0]
    lengths = sample_lengths(cm, b, 4000, np.random.default_rng(7))
    p, _ = chi2_pvalue(lengths, gamma[b])
    assert p > 1e-4
    shifted = np.roll(gamma[b], 1)
    shifted[0] = 0.0
    p2, _ = chi2_pvalue(lengths, shifted)
    assert p2 < 1e-6
