"""Tiny covariance models for brute-force tests, and a parse-tree enumerator.

make_cm_text(layout, seed) writes an INFERNAL1/a model whose guide tree is
`layout` (a list of node type names in preorder), with random probabilities.
The state layout follows Infernal's CM conventions (src/infernal.h:193-243;
states of a node in the order S/B/MP/ML/MR/D then IL/IR; a state's children
are the later inserts of its own node followed by the non-insert states of the
next node; B's children are the BEGL_S and BEGR_S states), as read by
read_asc_1p1_cm (src/cm_file.c:1539).

enumerate_scores(cm, codes, v, i, j) lists the total score of every parse tree
rooted at state v that emits codes[i..j] (1-based, inclusive), scoring a parse
as the sum of its transition, emission and local-end scores.  This is the
brute-force method of src/itest_brute.c applied to the scanners' root cells.
"""
import math

import numpy as np

NODE_STATES = {
    "ROOT": ["S", "IL", "IR"],
    "BIF": ["B"],
    "MATP": ["MP", "ML", "MR", "D", "IL", "IR"],
    "MATL": ["ML", "D", "IL"],
    "MATR": ["MR", "D", "IR"],
    "BEGL": ["S"],
    "BEGR": ["S", "IL"],
    "END": ["E"],
}
INSERTS = ("IL", "IR")


def _tree_children(layout):
    """Child node(s) of each node in a preorder layout."""
    n = len(layout)
    kids = [[] for _ in range(n)]

    def walk(nd):
        t = layout[nd]
        if t == "END":
            return nd + 1
        if t == "BIF":
            assert layout[nd + 1] == "BEGL"
            kids[nd].append(nd + 1)
            nxt = walk(nd + 1)
            assert layout[nxt] == "BEGR"
            kids[nd].append(nxt)
            return walk(nxt)
        kids[nd].append(nd + 1)
        return walk(nd + 1)

    end = walk(0)
    assert end == n, "layout is not a single preorder tree"
    return kids


def make_cm_text(layout, seed, W=8, elself=None, name="tiny"):
    rng = np.random.default_rng(seed)
    kids = _tree_children(layout)
    states = []  # (node, type)
    first = []
    for nd, t in enumerate(layout):
        first.append(len(states))
        for s in NODE_STATES[t]:
            states.append((nd, s))
    M = len(states)
    children = [[] for _ in range(M)]
    for v, (nd, s) in enumerate(states):
        t = layout[nd]
        if s == "E":
            continue
        if s == "B":
            children[v] = [first[kids[nd][0]], first[kids[nd][1]]]
            continue
        own = [u for u in range(first[nd], first[nd] + len(NODE_STATES[t])) if states[u][1] in INSERTS]
        if s in INSERTS:
            own = [u for u in own if u >= v]
        else:
            own = [u for u in own]
        nxt = kids[nd][0]
        split = [u for u in range(first[nxt], first[nxt] + len(NODE_STATES[layout[nxt]])) if states[u][1] not in INSERTS]
        children[v] = own + split
    parents = [[] for _ in range(M)]
    for v in range(M):
        if states[v][1] == "B":
            for c in children[v]:
                parents[c].append(v)
        else:
            for c in children[v]:
                parents[c].append(v)
    clen = sum(2 if t == "MATP" else 1 if t in ("MATL", "MATR") else 0 for t in layout)
    null = [0.25, 0.25, 0.25, 0.25]
    if elself is None:
        elself = -float(rng.uniform(0.05, 1.5))
    lines = ["INFERNAL1/a [tiny test model]", "NAME     %s" % name, "STATES   %d" % M, "NODES    %d" % len(layout),
             "CLEN     %d" % clen, "W        %d" % W, "ALPH     RNA", "RF       no", "CONS     no", "MAP      no",
             "PBEGIN   0.05", "PEND     0.05", "WBETA    1e-07", "QDBBETA1 1e-07", "QDBBETA2 1e-15",
             "N2OMEGA  1.52588e-05", "N3OMEGA  1.52588e-05", "ELSELF   %.6f" % elself,
             "NULL     " + " ".join("%.3f" % 0.0 for _ in null), "EFP7GF   -6.0 0.7", "CM"]
    for v, (nd, s) in enumerate(states):
        if v == first[nd]:
            lines.append("   [ %s %d ] - - - - - -" % (layout[nd], nd))
        ch = children[v]
        if s == "B":
            cf, cn = ch[0], ch[1]
        elif s == "E":
            cf, cn = -1, 0
        else:
            cf, cn = ch[0], len(ch)
        pl = max(parents[v]) if parents[v] else -1
        pn = len(parents[v])
        fields = ["%s" % s, str(v), str(pl), str(pn), str(cf), str(cn), "0", "0", str(W), str(W)]
        if s not in ("B", "E"):
            p = rng.dirichlet(np.ones(cn) * 1.5)
            p = np.maximum(p, 1e-3)
            p = p / p.sum()
            fields += ["%.6f" % math.log2(x) for x in p]
        if s in ("ML", "MR", "IL", "IR"):
            p = rng.dirichlet(np.ones(4) * 2)
            fields += ["%.6f" % math.log2(x / 0.25) for x in p]
        elif s == "MP":
            p = rng.dirichlet(np.ones(16) * 2)
            fields += ["%.6f" % math.log2(x / 0.0625) for x in p]
        lines.append(" ".join(fields))
    lines.append("//")
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------------------
# brute force

class Tables:
    """Configured score tables read back from the library's CM_t."""

    def __init__(self, cm):
        c = cm.c
        self.M = c.M
        self.sttype = [ord(c.sttype[v]) for v in range(c.M)]
        self.stid = [ord(c.stid[v]) for v in range(c.M)]
        self.cfirst = [c.cfirst[v] for v in range(c.M)]
        self.cnum = [c.cnum[v] for v in range(c.M)]
        self.tsc = [[c.tsc[v][k] for k in range(6)] for v in range(c.M)]
        self.endsc = [c.endsc[v] for v in range(c.M)]
        self.beginsc = [c.beginsc[v] for v in range(c.M)]
        self.el = c.el_selfsc
        self.local_begin = bool(c.flags & (1 << 10))
        self.esc = []
        for v in range(c.M):
            t = self.sttype[v]
            if t in (2, 3, 4, 5):
                self.esc.append([c.oesc[v][x] for x in range(18)])
            elif t == 1:
                self.esc.append([c.oesc[v][x] for x in range(18 * 18)])
            else:
                self.esc.append(None)


IMP = -1e30


def enumerate_scores(tb, x, v, i, j, memo=None, collapse_empty=False):
    """All parse scores for state v over x[i..j] (x is 1-based: x[0] unused).

    collapse_empty: replace the parses of an empty span (i > j) by the best of
    them.  The scanners initialise every d = 0 cell by max, also for Inside
    (src/cm_mx.c:6039-6138), so an Inside root cell sums over parse trees in
    which each empty subtree is fixed to its best completion."""
    if memo is None:
        memo = {}
    key = (v, i, j)
    if key in memo:
        return memo[key]
    out = _enumerate(tb, x, v, i, j, memo, collapse_empty)
    if collapse_empty and j < i and len(out) > 1:
        out = [max(out)]
    memo[key] = out
    return out


def _enumerate(tb, x, v, i, j, memo, collapse_empty):
    st = tb.sttype[v]
    d = j - i + 1
    out = []
    rec = lambda u, a, b: enumerate_scores(tb, x, u, a, b, memo, collapse_empty)
    if st == 7:  # E
        if d == 0:
            out.append(0.0)
        return out
    if st == 8:  # B
        if tb.endsc[v] > IMP:
            out.append(tb.endsc[v] + tb.el * d)
        w, y = tb.cfirst[v], tb.cnum[v]
        for k in range(i - 1, j + 1):
            left = rec(w, i, k)
            right = rec(y, k + 1, j)
            out.extend(a + b for a in left for b in right)
        return out
    if st in (2, 4):      # ML, IL
        if d < 1:
            return out
        e, ii, jj = tb.esc[v][x[i]], i + 1, j
    elif st in (3, 5):    # MR, IR
        if d < 1:
            return out
        e, ii, jj = tb.esc[v][x[j]], i, j - 1
    elif st == 1:         # MP
        if d < 2:
            return out
        e, ii, jj = tb.esc[v][x[i] * 18 + x[j]], i + 1, j - 1
    else:                 # D, S
        e, ii, jj = 0.0, i, j
    if e <= IMP:
        return out
    if v != 0 and tb.endsc[v] > IMP:
        out.append(e + tb.endsc[v] + tb.el * (jj - ii + 1))
    for k in range(tb.cnum[v]):
        t = tb.tsc[v][k]
        if t <= IMP:
            continue
        for s in rec(tb.cfirst[v] + k, ii, jj):
            out.append(e + t + s)
    return out


def log2sumexp2(xs):
    m = max(xs)
    return m + math.log2(sum(2.0 ** (a - m) for a in xs))


def root_score(tb, x, i, j, inside):
    """The scanner's root cell for x[i..j]: root parses combined (max or sum),
    then local begins taken by max (src/cm_dpsearch.c:562-616, :2200-2248)."""
    memo = {}
    cand = []
    root = enumerate_scores(tb, x, 0, i, j, memo, collapse_empty=inside)
    best = None
    if root:
        best = log2sumexp2(root) if inside else max(root)
    if tb.local_begin:
        for y in range(1, tb.M):
            if tb.beginsc[y] <= IMP:
                continue
            s = enumerate_scores(tb, x, y, i, j, memo, collapse_empty=inside)
            if not s:
                continue
            val = tb.beginsc[y] + (log2sumexp2(s) if inside else max(s))
            cand.append(val)
    if cand:
        m = max(cand)
        if best is None or m > best:
            best = m
    return best
