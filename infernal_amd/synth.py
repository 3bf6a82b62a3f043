"""Seeded synthetic targets for the scanners: background sequence and
homologs sampled from a covariance model.

* markov2_background(n, gc, seed): iid-free background from a seeded
  second-order Markov chain whose stationary GC content is `gc`.  Generated as
  many independent chains of `seg` residues advanced together, so 100 Mb takes
  seconds in numpy.
* emit(cm, rng): one sequence sampled from the model's global parse-tree
  distribution (transition and emission probabilities), as cmemit does with
  EmitParsetree (src/cm_parsetree.c:1890): start at ROOT_S, emit by state
  type, pick a child by t[v][], descend both branches of a bifurcation.
* plant(background, homologs, seed): overwrite seeded, non-overlapping
  positions of the background with the homolog fragments.
"""
import bisect

import numpy as np

from . import B_st, D_st, E_st, IL_st, IR_st, ML_st, MP_st, MR_st, S_st


def markov2_background(n, gc=0.40, seed=1, seg=10000, sample_seed=None):
    """n residues of a second-order Markov chain with stationary GC content
    gc; the chain's conditionals are drawn from `seed`, the sequence from
    the same stream (or from sample_seed: another sample of the same chain)."""
    rng = np.random.default_rng(seed)
    base = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    # per-context conditionals, then rescale them so the chain's stationary
    # composition matches `base` (a few fixed-point rounds)
    P = np.array([rng.dirichlet(base * 40.0) for _ in range(16)])
    for _ in range(50):
        T = np.zeros((16, 16))
        for c in range(16):
            for x in range(4):
                T[c, (c % 4) * 4 + x] = P[c, x]
        w, V = np.linalg.eig(T.T)
        pi = np.real(V[:, np.argmin(np.abs(w - 1))])
        pi = pi / pi.sum()
        comp = np.zeros(4)
        for c in range(16):
            comp[c % 4] += pi[c]
        P = P * (base / comp)[None, :]
        P = P / P.sum(axis=1, keepdims=True)
    cdf = np.cumsum(P, axis=1)
    if sample_seed is not None:
        rng = np.random.default_rng(sample_seed)
    nseg = (n + seg - 1) // seg
    out = np.empty((nseg, seg), dtype=np.uint8)
    x1 = rng.choice(4, size=nseg, p=base)
    x2 = rng.choice(4, size=nseg, p=base)
    out[:, 0] = x1
    out[:, 1] = x2
    u = rng.random((seg, nseg), dtype=np.float32)
    for t in range(2, seg):
        ctx = out[:, t - 2].astype(np.int64) * 4 + out[:, t - 1]
        out[:, t] = (u[t][:, None] > cdf[ctx]).sum(axis=1).clip(0, 3)
    return out.reshape(-1)[:n].copy()


def _emit_tables(cm):
    """Per-state sampling tables, read from the model once and cached on the
    Python CM object: state types, children, and for every state the
    cumulative distributions of its transitions and of its emissions."""
    tab = getattr(cm, "_emit_tab", None)
    if tab is not None:
        return tab
    c = cm.c
    M = c.M
    st = [ord(c.sttype[v]) for v in range(M)]
    cfirst = [c.cfirst[v] for v in range(M)]
    cnum = [c.cnum[v] for v in range(M)]

    def cdf(p):
        # the cumulative sums numpy's Generator.choice(len(p), p=p/sum(p))
        # draws against, so that sampling by bisection below picks the same
        # index for the same uniform draw
        p = np.asarray(p, dtype=np.float64)
        p = p / p.sum()
        q = p.cumsum()
        q /= q[-1]
        return q.tolist()

    tcdf, ecdf = [None] * M, [None] * M
    for v in range(M):
        if st[v] in (E_st, B_st):
            continue
        probs = ([c.root_trans[k] for k in range(cnum[0])] if c.root_trans else [c.t[0][k] for k in range(cnum[0])]) \
            if v == 0 else [c.t[v][k] for k in range(cnum[v])]
        tcdf[v] = cdf(probs)
        if st[v] in (ML_st, IL_st, MR_st, IR_st):
            ecdf[v] = cdf([c.e[v][k] for k in range(4)])
        elif st[v] == MP_st:
            ecdf[v] = cdf([c.e[v][k] for k in range(16)])
    tab = (st, cfirst, cnum, tcdf, ecdf)
    cm._emit_tab = tab
    return tab


def emit(cm, rng, max_len=100000):
    """Sample one sequence (residue codes 0..3) from the model (global mode).
    Each choice takes one uniform draw from rng, as Generator.choice with
    probabilities does."""
    st, cfirst, cnum, tcdf, ecdf = _emit_tables(cm)
    rand = rng.random

    def pick(q):
        return bisect.bisect_right(q, rand())

    # iterative walk; a node emits left residues on the way down and right
    # residues on the way back up
    out = []
    stack = [("v", 0)]
    while stack:
        kind, v = stack.pop()
        if kind == "r":          # deferred right emission
            out.append(v)
            continue
        s = st[v]
        if s == E_st:
            continue
        if s == B_st:
            stack.append(("v", cnum[v]))     # right subtree after the left one
            stack.append(("v", cfirst[v]))
            continue
        if s in (ML_st, IL_st):
            out.append(pick(ecdf[v]))
        elif s in (MR_st, IR_st):
            stack.append(("r", pick(ecdf[v])))
        elif s == MP_st:
            xy = pick(ecdf[v])
            out.append(xy // 4)
            stack.append(("r", xy % 4))
        k = pick(tcdf[v])
        stack.append(("v", cfirst[v] + k))
        if len(out) > max_len:
            break
    return np.array(out, dtype=np.uint8)


def homologs(cm, n, seed, frac=(0.5, 1.0)):
    """n emitted sequences, each cut to a random fragment of `frac` of its length."""
    rng = np.random.default_rng(seed)
    res = []
    for _ in range(n):
        s = emit(cm, rng)
        f = rng.uniform(*frac)
        L = max(1, int(round(len(s) * f)))
        a = int(rng.integers(0, len(s) - L + 1))
        res.append(s[a:a + L])
    return res


def plant(background, frags, seed):
    """Overwrite non-overlapping seeded positions; returns (codes, starts)."""
    rng = np.random.default_rng(seed)
    seq = background.copy()
    n = len(seq)
    total = sum(len(f) for f in frags)
    if total >= n:
        raise ValueError("homologs do not fit in the background")
    # spread: pick sorted gap sizes that sum to n - total
    gaps = rng.multinomial(n - total, np.ones(len(frags) + 1) / (len(frags) + 1))
    starts = []
    pos = 0
    for g, f in zip(gaps[:-1], frags):
        pos += int(g)
        seq[pos:pos + len(f)] = f
        starts.append(pos + 1)
        pos += len(f)
    return seq, starts


def truncated_windows(cm, n, seed, lmin=200, lmax=2000, min_overlap=100, pool=100):
    """n windows of uniform length in [lmin, lmax], half holding an emitted
    homolog clipped by the window at a random offset (so it is cut at the
    window's 5' end, its 3' end, both or neither), half uniform random ACGU
    decoys; concatenated into one target.  The homologs are drawn from a pool
    of `pool` emitted sequences (emission is the slow part).  Returns
    (codes, windows, is_hom) with 1-based inclusive (i, j) window bounds."""
    rng = np.random.default_rng(seed)
    seqs = [emit(cm, rng) for _ in range(max(1, min(pool, (n + 1) // 2)))]
    parts, wins, hom = [], [], []
    pos = 1
    for k in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        w = rng.integers(0, 4, L).astype(np.uint8)
        if k % 2 == 0:
            s = seqs[int(rng.integers(0, len(seqs)))]
            ov = min(min_overlap, len(s), L)
            o = int(rng.integers(-(len(s) - ov), L - ov + 1))   # s starts at window offset o
            a, b = max(0, -o), min(len(s), L - o)
            w[o + a:o + b] = s[a:b]
        parts.append(w)
        wins.append((pos, pos + L - 1))
        hom.append(k % 2 == 0)
        pos += L
    return np.concatenate(parts), wins, np.array(hom)
