"""Covariance-model construction from an alignment and its consensus
structure, with cmbuild's default options (a restatement of the parts of
src/cmbuild.c, src/cm_modelmaker.c, src/cm_parsetree.c, src/eweight.c and
src/prior.c on that path), used to make the benchmark and test models.

  guide_tree(ct, mat)  HandModelmaker's guide tree (src/cm_modelmaker.c:203-425)
  states(root)         cm_from_guide (src/cm_modelmaker.c:596-870)
  rebalance(layout)    CMRebalance state numbering (src/cm.c:1821-1990)
  pb_weights           Easel's position-based weights as cmbuild calls them
  match_columns        AssignMatchColumnsForMsa (src/cm_modelmaker.c:511-560)
  count_parse          Transmogrify + ParsetreeCountExceptTruncatedMPs
  entropy_weight       cm_EntropyWeight (src/eweight.c:120-230)
  priorify             PriorifyCM with the Dirichlet mixtures of a prior file
  detach_dual_inserts  cm_find_and_detach_dual_inserts (src/cm_modelmaker.c)
  build_cm_from_alignment / build_cm   the whole path, bands and W from
                       infx_CalculateQueryDependentBands, INFERNAL1/a text

Rebuilding the reference's tutorial models (tRNA5, 5S_rRNA, Cobalamin) from
their alignments reproduces the shipped files: same EFFN, W and bands, every
score within print rounding.  Not covered: --hand, truncated ('~')
sequences, non-default weighting, prior or entropy options.
"""
import math

import numpy as np

# node types (src/infernal.h:208-216)
BIF_nd, MATP_nd, MATL_nd, MATR_nd, BEGL_nd, BEGR_nd, ROOT_nd, END_nd = range(8)
NODE_NAMES = {BIF_nd: "BIF", MATP_nd: "MATP", MATL_nd: "MATL", MATR_nd: "MATR", BEGL_nd: "BEGL", BEGR_nd: "BEGR",
              ROOT_nd: "ROOT", END_nd: "END"}
# state types (src/infernal.h:190-200)
D_st, MP_st, ML_st, MR_st, IL_st, IR_st, S_st, E_st, B_st = 0, 1, 2, 3, 4, 5, 6, 7, 8
STATE_NAMES = {D_st: "D", MP_st: "MP", ML_st: "ML", MR_st: "MR", IL_st: "IL", IR_st: "IR", S_st: "S", E_st: "E", B_st: "B"}

# connectable states in a child / parent node (src/cm_modelmaker.c:612-614)
CHILD_COUNT = {BIF_nd: 1, MATP_nd: 4, MATL_nd: 2, MATR_nd: 2, BEGL_nd: 1, BEGR_nd: 1, ROOT_nd: 0, END_nd: 1}
PARENT_COUNT = {BIF_nd: 1, MATP_nd: 6, MATL_nd: 3, MATR_nd: 3, BEGL_nd: 1, BEGR_nd: 2, ROOT_nd: 3, END_nd: 0}

_PAIR_OPEN = "<([{"
_PAIR_CLOSE = ">)]}"


def ss_to_ct(ss):
    """WUSS/dot-bracket consensus structure -> ct[1..n] (0 = unpaired);
    pseudoknot letters are treated as unpaired."""
    n = len(ss)
    ct = [0] * (n + 1)
    stacks = {c: [] for c in _PAIR_OPEN}
    for p, ch in enumerate(ss, start=1):
        if ch in _PAIR_OPEN:
            stacks[ch].append(p)
        elif ch in _PAIR_CLOSE:
            o = _PAIR_OPEN[_PAIR_CLOSE.index(ch)]
            if not stacks[o]:
                raise ValueError("unbalanced structure at %d" % p)
            q = stacks[o].pop()
            ct[p], ct[q] = q, p
    if any(stacks[c] for c in stacks):
        raise ValueError("unbalanced structure")
    return ct


class Node:
    __slots__ = ("type", "i", "j", "left", "right")

    def __init__(self, t, i, j):
        self.type, self.i, self.j = t, i, j
        self.left = self.right = None


def guide_tree(ct, mat=None):
    """Guide tree (HandModelmaker, src/cm_modelmaker.c:203-425) over alignment
    columns 1..n with consensus pairing ct (0 = unpaired or non-consensus) and
    mat[c] True for consensus columns (None: every column).  Per subsequence
    i..j (both consensus columns): END if empty, MATL if i unpaired, MATR if j
    unpaired, MATP if i pairs j, else BIF split at the k (stepping over closed
    stems from ct[i]+1) whose two sides have the closest consensus lengths,
    BEGL i..k-1 and BEGR k..j.  Non-consensus columns are stepped over; they
    are the insert columns of the node before them.  Node spans are alignment
    columns.  Returns the ROOT node."""
    n = len(ct) - 1
    if mat is None:
        mat = [True] * (n + 2)
    a2c = [0] * (n + 2)
    c = 0
    for col in range(1, n + 1):
        if mat[col]:
            c += 1
            a2c[col] = c

    def next_mat(i, j):
        while i <= j and not mat[i]:
            i += 1
        return i

    def prev_mat(j, i):
        while j >= i and not mat[j]:
            j -= 1
        return j

    def dummy(i, j):
        if i > j:
            return Node(END_nd, i, j)
        if ct[i] == 0:
            nd = Node(MATL_nd, i, j)
            nd.left = dummy(next_mat(i + 1, j), j)
            return nd
        if ct[j] == 0:
            nd = Node(MATR_nd, i, j)
            nd.left = dummy(i, prev_mat(j - 1, i))
            return nd
        if ct[i] == j:
            nd = Node(MATP_nd, i, j)
            i2 = next_mat(i + 1, j)
            nd.left = dummy(i2, prev_mat(j - 1, i2))
            return nd
        # bifurcation: the most balanced split point (first one on ties)
        nd = Node(BIF_nd, i, j)
        bestk, bestdiff = ct[i] + 1, n + 2
        k = ct[i] + 1
        while k <= ct[j]:
            kp = k
            while not mat[kp]:
                kp += 1
            diff = abs((a2c[kp] - a2c[i]) - (a2c[j] - a2c[kp] + 1))
            if diff < bestdiff:
                bestdiff, bestk = diff, k
            while ct[k] == 0:
                k += 1
            k = ct[k] + 1
        bl = Node(BEGL_nd, i, bestk - 1)
        bl.left = dummy(i, bestk - 1)
        br = Node(BEGR_nd, bestk, j)
        br.left = dummy(next_mat(bestk, j), j)
        nd.left, nd.right = bl, br
        return nd

    root = Node(ROOT_nd, 1, n)
    i = next_mat(1, n)
    root.left = dummy(i, prev_mat(n, i))
    return root


def preorder(root):
    out, stack = [], [(root, None)]
    while stack:
        nd, parent = stack.pop()
        out.append((nd, parent))
        if nd.type == BIF_nd:
            stack.append((nd.right, nd))
            stack.append((nd.left, nd))
        elif nd.left is not None:
            stack.append((nd.left, nd))
    return out


def states(root):
    """cm_from_guide: per node its states; per state (type, node, cfirst, cnum,
    plast, pnum).  Returns (node_types, nodemap, sttype, ndidx, cfirst, cnum,
    plast, pnum)."""
    order = preorder(root)
    nd_types, nodemap = [], []
    st, ndidx, cfirst, cnum, plast, pnum = [], [], [], [], [], []
    bif_of = {}

    def add(t, node, cf, cn, pl, pn):
        st.append(t); ndidx.append(node); cfirst.append(cf); cnum.append(cn); plast.append(pl); pnum.append(pn)

    for node, (nd, parent) in enumerate(order):
        s = len(st)
        nd_types.append(nd.type)
        nodemap.append(s)
        nxt = nd.left.type if nd.left is not None else None
        prv = parent.type if parent is not None else None
        if nd.type == BIF_nd:
            add(B_st, node, s + 1, -1, s - 1, PARENT_COUNT[prv])
            bif_of[id(nd.right)] = s
        elif nd.type == MATP_nd:
            cc = CHILD_COUNT[nxt]
            add(MP_st, node, s + 4, 2 + cc, s - 1, PARENT_COUNT[prv])
            add(ML_st, node, s + 4, 2 + cc, s - 1, PARENT_COUNT[prv])
            add(MR_st, node, s + 4, 2 + cc, s - 1, PARENT_COUNT[prv])
            add(D_st, node, s + 4, 2 + cc, s - 1, PARENT_COUNT[prv])
            add(IL_st, node, s + 4, 2 + cc, s + 4, 5)
            add(IR_st, node, s + 5, 1 + cc, s + 5, 6)
        elif nd.type in (MATL_nd, MATR_nd):
            cc = CHILD_COUNT[nxt]
            add(ML_st if nd.type == MATL_nd else MR_st, node, s + 2, 1 + cc, s - 1, PARENT_COUNT[prv])
            add(D_st, node, s + 2, 1 + cc, s - 1, PARENT_COUNT[prv])
            add(IL_st if nd.type == MATL_nd else IR_st, node, s + 2, 1 + cc, s + 2, 3)
        elif nd.type == BEGL_nd:
            add(S_st, node, s + 1, CHILD_COUNT[nxt], s - 1, 1)
        elif nd.type == BEGR_nd:
            b = bif_of[id(nd)]
            cnum[b] = s
            cc = CHILD_COUNT[nxt]
            add(S_st, node, s + 1, 1 + cc, b, 1)
            add(IL_st, node, s + 1, 1 + cc, s + 1, 2)
        elif nd.type == ROOT_nd:
            cc = CHILD_COUNT[nxt]
            add(S_st, node, s + 1, 2 + cc, -1, 0)
            add(IL_st, node, s + 1, 2 + cc, s + 1, 2)
            add(IR_st, node, s + 2, 1 + cc, s + 2, 3)
        elif nd.type == END_nd:
            add(E_st, node, -1, 0, s - 1, PARENT_COUNT[prv])
    return nd_types, nodemap, st, ndidx, cfirst, cnum, plast, pnum


def rebalance(layout):
    """CMRebalance (src/cm.c:1821-1990): renumber the states in a modified
    preorder that visits, at each bifurcation, the child subgraph needing fewer
    extra CYK decks first (weight: E 1, B 1 + min(children), others as the next
    state).  Node numbering is unchanged.  layout = states() output; returns the
    same tuple renumbered, plus newidx (old state -> new state)."""
    nd_types, nodemap, st, ndidx, cfirst, cnum, plast, pnum = layout
    M = len(st)
    wgt = [0] * M
    for v in range(M - 1, -1, -1):
        if st[v] == E_st:
            wgt[v] = 1
        elif st[v] == B_st:
            wgt[v] = 1 + min(wgt[cfirst[v]], wgt[cnum[v]])
        else:
            wgt[v] = wgt[v + 1]
    n_st, n_nd, n_cf, n_cn, n_pl, n_pn = [0] * M, [0] * M, [0] * M, [0] * M, [0] * M, [0] * M
    newidx = [0] * M
    pda = []
    v, z = 0, M - 1
    for nv in range(M):
        newidx[v] = nv
        n_st[nv], n_nd[nv], n_pn[nv] = st[v], ndidx[v], pnum[v]
        n_pl[nv] = newidx[plast[v]] if nv != 0 else -1
        if st[v] == B_st:
            w, y = cfirst[v], cnum[v]
            if wgt[w] <= wgt[y]:
                pda += [y, z]
                v, z = w, y - 1
                n_cf[nv], n_cn[nv] = nv + 1, nv + y - w + 1
            else:
                pda += [w, y - 1]
                n_cf[nv], n_cn[nv] = nv + z - y + 2, nv + 1
                v = y
        elif st[v] == E_st:
            n_cf[nv], n_cn[nv] = -1, 0
            if pda:
                z = pda.pop()
                v = pda.pop()
        else:
            n_cf[nv], n_cn[nv] = nv + (cfirst[v] - v), cnum[v]
            v += 1
    n_map = [newidx[s] for s in nodemap]
    return (nd_types, n_map, n_st, n_nd, n_cf, n_cn, n_pl, n_pn), newidx


def layout_for_structure(ct):
    """Guide tree, states and the rebalanced numbering cmbuild writes."""
    return rebalance(states(guide_tree(ct)))[0]


# ---------------------------------------------------------------------------
# parameters

STID_NAMES = {
    (ROOT_nd, S_st): "ROOT_S", (ROOT_nd, IL_st): "ROOT_IL", (ROOT_nd, IR_st): "ROOT_IR",
    (BEGL_nd, S_st): "BEGL_S", (BEGR_nd, S_st): "BEGR_S", (BEGR_nd, IL_st): "BEGR_IL",
    (MATP_nd, MP_st): "MATP_MP", (MATP_nd, ML_st): "MATP_ML", (MATP_nd, MR_st): "MATP_MR", (MATP_nd, D_st): "MATP_D",
    (MATP_nd, IL_st): "MATP_IL", (MATP_nd, IR_st): "MATP_IR",
    (MATL_nd, ML_st): "MATL_ML", (MATL_nd, D_st): "MATL_D", (MATL_nd, IL_st): "MATL_IL",
    (MATR_nd, MR_st): "MATR_MR", (MATR_nd, D_st): "MATR_D", (MATR_nd, IR_st): "MATR_IR",
}


def read_prior(path):
    """A Dirichlet-mixture prior file in cmbuild's format (Prior_Read,
    src/prior.c:84-185): 'Dirichlet', the number of transition sets, then per
    set '<unique state> <next node type>' and a mixture, then the mixtures of
    consensus pairs, consensus singlets and non-consensus singlets.  A mixture
    is K, Q, then per component its weight and K alphas."""
    toks = []
    for line in open(path):
        line = line.split("#", 1)[0]
        toks += line.split()
    pos = [0]

    def nxt():
        pos[0] += 1
        return toks[pos[0] - 1]

    def mixture():
        K, Q = int(nxt()), int(nxt())
        pq, alpha = [], []
        for _ in range(Q):
            pq.append(float(nxt()))
            alpha.append([float(nxt()) for _ in range(K)])
        return np.array(pq), np.array(alpha)

    if nxt().lower() != "dirichlet":
        raise ValueError("not a Dirichlet prior file")
    nsets = int(nxt())
    tsets = {}
    for _ in range(nsets):
        sname, nname = nxt(), nxt()
        tsets[(sname, nname)] = mixture()
    return {"t": tsets, "mbp": mixture(), "mnt": mixture(), "i": mixture()}


def mixture_mean(mix, counts):
    """Posterior mean parameters of a Dirichlet mixture given counts
    (esl_mixdchlet_MPParameters): component posteriors from the Dirichlet-
    multinomial likelihood, then the mixture of the component posterior means."""
    pq, alpha = mix
    c = np.asarray(counts, dtype=np.float64)
    csum = c.sum()
    logp = np.full(len(pq), -np.inf)
    for q in range(len(pq)):
        if pq[q] > 0:
            a = alpha[q]
            logp[q] = (math.log(pq[q]) + math.lgamma(a.sum()) - math.lgamma(a.sum() + csum)
                       + sum(math.lgamma(a[i] + c[i]) - math.lgamma(a[i]) for i in range(len(a))))
    w = np.exp(logp - logp.max())
    w /= w.sum()
    p = np.zeros(len(c))
    for q in range(len(pq)):
        p += w[q] * (c + alpha[q]) / (csum + alpha[q].sum())
    return p / p.sum()


class Model:
    """A CM under construction: layout (cmbuild numbering) plus float64
    transition counts/probabilities t[M][6] and emissions e[M][16]."""

    def __init__(self, ct):
        self.ct = ct
        self.clen = len(ct) - 1
        self.root = guide_tree(ct)
        (self.nd_types, self.nodemap, self.st, self.ndidx, self.cfirst, self.cnum, self.plast, self.pnum) = \
            layout_for_structure(ct)
        # node spans in consensus coordinates (preorder = node index)
        self.spans = [(nd.i, nd.j) for nd, _ in preorder(self.root)]
        self.M = len(self.st)
        self.t = np.zeros((self.M, 6))
        self.e = np.zeros((self.M, 16))

    def stid(self, v):
        return STID_NAMES.get((self.nd_types[self.ndidx[v]], self.st[v]))

    def main_state(self, nd):
        return self.nodemap[nd]


# digital alignment codes: 0-3 ACGU, GAP, degenerate residues, MISSING ('~')
GAP = 4
DEGEN = "RYMKSWHBVDN"
MISSING = 16
_DEGEN_SETS = {"R": "AG", "Y": "CU", "M": "AC", "K": "GU", "S": "CG", "W": "AU", "H": "ACU", "B": "CGU",
               "V": "ACG", "D": "AGU", "N": "ACGU"}
# fractional canonical counts of each code (Easel's esl_abc_FCount with weight 1)
FCOUNT = np.zeros((17, 4))
for _x in range(4):
    FCOUNT[_x, _x] = 1.0
for _k, _ch in enumerate(DEGEN):
    for _r in _DEGEN_SETS[_ch]:
        FCOUNT[5 + _k, "ACGU".index(_r)] = 1.0 / len(_DEGEN_SETS[_ch])


def read_stockholm(path):
    """Aligned sequences (in file order, blocks joined) and #=GC lines of a
    one-alignment Stockholm file."""
    seqs, order, gc = {}, [], {}
    with open(path) as fh:
        for line in fh:
            if not line.strip() or line.startswith("# STOCKHOLM") or line.startswith("//"):
                continue
            if line.startswith("#=GC"):
                _, tag, val = line.split(None, 2)
                gc[tag] = gc.get(tag, "") + val.strip()
                continue
            if line.startswith("#"):
                continue
            name, seq = line.split()
            if name not in seqs:
                seqs[name] = ""
                order.append(name)
            seqs[name] += seq
    return [seqs[n] for n in order], gc


def digitize_alignment(aseqs):
    """Aligned text sequences -> int array of the codes above (T reads as U;
    '.', '-' and '_' are gaps; unknown letters read as N)."""
    lut = {"A": 0, "C": 1, "G": 2, "U": 3, "T": 3, ".": GAP, "-": GAP, "_": GAP, "~": MISSING}
    for k, ch in enumerate(DEGEN):
        lut[ch] = 5 + k
    out = np.zeros((len(aseqs), len(aseqs[0])), dtype=np.int64)
    for r, seq in enumerate(aseqs):
        out[r] = [lut.get(ch, 5 + DEGEN.index("N")) for ch in seq.upper()]
    return out


def pb_weights(ax, symfrac=0.5):
    """Henikoff position-based weights as cmbuild computes them by default
    (Easel's esl_msaweight_PB_adv, ignoring RF): over the columns whose
    residue fraction (residues / (residues + gaps)) is >= symfrac, a sequence
    with canonical residue a in a column with r canonical residue types and
    n_a copies of a gets 1/(r n_a); each sequence's sum is divided by its
    number of canonical residues in those columns; the weights are scaled to
    sum to the number of sequences.  (Pinned by the effective sequence
    numbers and match emissions of the tutorial models.)"""
    nseq, alen = ax.shape
    if nseq == 1:
        return np.ones(1)
    isres = (ax < GAP) | ((ax > GAP) & (ax < MISSING))
    denom = isres.sum(0) + (ax == GAP).sum(0)
    cols = [c for c in range(alen) if denom[c] > 0 and isres[:, c].sum() / denom[c] >= symfrac]
    w = np.zeros(nseq)
    for c in cols:
        col = ax[:, c]
        canon = col < GAP
        cnt = np.bincount(col[canon], minlength=4)
        ntypes = int((cnt > 0).sum())
        if ntypes == 0:
            continue
        w[canon] += 1.0 / (ntypes * cnt[col[canon]])
    rlen = (ax[:, cols] < GAP).sum(1)
    w = np.where(rlen > 0, w / np.maximum(rlen, 1), w)
    return w * (nseq / w.sum())


def henikoff_weights(aln):
    """pb_weights of an ungapped alignment of canonical residues."""
    return pb_weights(np.asarray(aln))


def match_columns(ax, wts, symfrac=0.5):
    """AssignMatchColumnsForMsa (src/cm_modelmaker.c:511-560), weighted: a
    column is consensus if its weighted residue fraction is >= symfrac.
    Returns mat[0..alen+1] (1-based, False at both ends)."""
    isres = (ax < GAP) | ((ax > GAP) & (ax < MISSING))
    w = np.asarray(wts, dtype=np.float32)
    mat = [False] * (ax.shape[1] + 2)
    for c in range(ax.shape[1]):
        r = np.float32(w[isres[:, c]].sum())
        tot = np.float32(w[isres[:, c] | (ax[:, c] == GAP)].sum())
        mat[c + 1] = bool(r > 0 and r / tot >= symfrac)
    return mat


# state offsets within a node, in cm_from_guide order
_OFF = {(ROOT_nd, S_st): 0, (ROOT_nd, IL_st): 1, (ROOT_nd, IR_st): 2, (BIF_nd, B_st): 0,
        (MATP_nd, MP_st): 0, (MATP_nd, ML_st): 1, (MATP_nd, MR_st): 2, (MATP_nd, D_st): 3, (MATP_nd, IL_st): 4,
        (MATP_nd, IR_st): 5, (MATL_nd, ML_st): 0, (MATL_nd, D_st): 1, (MATL_nd, IL_st): 2,
        (MATR_nd, MR_st): 0, (MATR_nd, D_st): 1, (MATR_nd, IR_st): 2, (BEGL_nd, S_st): 0, (BEGR_nd, S_st): 0,
        (BEGR_nd, IL_st): 1, (END_nd, E_st): 0}


def count_parse(m, gorder, row, wgt, t, e):
    """Transmogrify + ParsetreeCountExceptTruncatedMPs
    (src/cm_modelmaker.c:973-1400, src/cm_parsetree.c:301-365) for one aligned
    sequence without missing data: walk the guide tree in preorder; each
    consensus node uses its match, split or delete state by the gap pattern
    of its columns, and each residue in its insert columns adds an IL (left
    inserts, left to right) or IR (right inserts, right to left) step.  Adds
    wgt to every transition along the parse (B and E have none) and the
    residues' fractional counts to the emitting states."""
    w = float(np.float32(wgt))
    isgap = row == GAP
    isres = ~isgap & (row != MISSING)
    node_index = {id(nd): k for k, (nd, _) in enumerate(gorder)}
    stack = []
    cur = [None]          # state of the trace node the next one attaches to

    def step(node, sttype, left=None, right=None):
        z = m.nodemap[node] + _OFF[(m.nd_types[node], sttype)]
        v = cur[0]
        if v is not None and m.st[v] not in (B_st, E_st):
            t[v][z - m.cfirst[v]] += w
        if sttype == MP_st:
            e[z] += np.outer(FCOUNT[row[left - 1]], FCOUNT[row[right - 1]]).ravel() * w
        elif sttype in (ML_st, IL_st):
            e[z, :4] += FCOUNT[row[left - 1]] * w
        elif sttype in (MR_st, IR_st):
            e[z, :4] += FCOUNT[row[right - 1]] * w
        cur[0] = z

    for node, (nd, _) in enumerate(gorder):
        typ = nd.type
        nx = nd.left
        if typ == ROOT_nd:
            cur[0] = None
            step(node, S_st)
            for i in range(nd.i, nx.i):
                if isres[i - 1]:
                    step(node, IL_st, left=i)
            for j in range(nd.j, nx.j, -1):
                if isres[j - 1]:
                    step(node, IR_st, right=j)
        elif typ == BIF_nd:
            step(node, B_st)
            stack.append(cur[0])
        elif typ == BEGL_nd:
            step(node, S_st)
        elif typ == BEGR_nd:
            cur[0] = None      # right child of the B: no transition counted
            stack.pop()
            step(node, S_st)
            for i in range(nd.i, nx.i):
                if isres[i - 1]:
                    step(node, IL_st, left=i)
        elif typ == END_nd:
            step(node, E_st)
        elif typ == MATP_nd:
            gl, gr = isgap[nd.i - 1], isgap[nd.j - 1]
            step(node, D_st if (gl and gr) else MR_st if gl else ML_st if gr else MP_st, left=nd.i, right=nd.j)
            for i in range(nd.i + 1, nx.i):
                if isres[i - 1]:
                    step(node, IL_st, left=i)
            for j in range(nd.j - 1, nx.j, -1):
                if isres[j - 1]:
                    step(node, IR_st, right=j)
        elif typ == MATL_nd:
            step(node, D_st if isgap[nd.i - 1] else ML_st, left=nd.i)
            for i in range(nd.i + 1, nx.i):
                if isres[i - 1]:
                    step(node, IL_st, left=i)
        elif typ == MATR_nd:
            step(node, D_st if isgap[nd.j - 1] else MR_st, right=nd.j)
            for j in range(nd.j - 1, nx.j, -1):
                if isres[j - 1]:
                    step(node, IR_st, right=j)


def zero_flanking_insert_counts(m, t):
    """cm_zero_flanking_insert_counts (src/cm_modelmaker.c:2171-2230): counts
    through ROOT_IL/ROOT_IR are moved onto ROOT_S's transitions into node 1,
    and the two inserts keep no counts."""
    t[0][0] = t[0][1] = 0.0
    for k in range(m.cnum[0] - 2):
        t[0][2 + k] += t[1][2 + k] + t[2][1 + k]
    t[1, :] = 0.0
    t[2, :] = 0.0


def priorify(m, t_counts, e_counts, pri):
    """PriorifyCM (src/prior.c:190-265): transition vectors by the set of
    (unique state, type of the node of the last child); MP pairs, consensus
    ML/MR singlets and non-consensus singlets (inserts, MATP ML/MR) by their
    emission mixtures."""
    t = np.zeros_like(t_counts)
    e = np.zeros_like(e_counts)
    for v in range(m.M):
        s = m.st[v]
        if s not in (B_st, E_st):
            last = m.cfirst[v] + m.cnum[v] - 1
            nxt = NODE_NAMES[m.nd_types[m.ndidx[last]]]
            mix = pri["t"][(m.stid(v), nxt)]
            t[v, :m.cnum[v]] = mixture_mean(mix, t_counts[v, :m.cnum[v]])
        sid = m.stid(v)
        if s == MP_st:
            e[v] = mixture_mean(pri["mbp"], e_counts[v])
        elif sid in ("MATL_ML", "MATR_MR"):
            e[v, :4] = mixture_mean(pri["mnt"], e_counts[v, :4])
        elif s in (IL_st, IR_st) or sid in ("MATP_ML", "MATP_MR"):
            e[v, :4] = mixture_mean(pri["i"], e_counts[v, :4])
    return t, e


def mean_match_relent(m, e):
    """cm_MeanMatchRelativeEntropy (src/eweight.c:378-440): relative entropy
    of MP pairs (vs the uniform pair background) and ML/MR singlets of MATL/MATR,
    summed and divided by the consensus length."""
    kl = 0.0
    for v in range(m.M):
        sid = m.stid(v)
        if sid == "MATP_MP":
            p = e[v]
            kl += float(np.sum(np.where(p > 0, p * np.log2(np.where(p > 0, p, 1) / (1 / 16)), 0)))
        elif sid in ("MATL_ML", "MATR_MR"):
            p = e[v, :4]
            kl += float(np.sum(np.where(p > 0, p * np.log2(np.where(p > 0, p, 1) / 0.25), 0)))
    return kl / m.clen


def entropy_weight(m, t_counts, e_counts, nseq, pri, etarget, min_neff=0.1):
    """cm_EntropyWeight (src/eweight.c:120-230): the effective sequence number
    whose rescaled, priorified model has mean match relative entropy etarget
    (bisection on [0, nseq] to 0.01; min_neff if even that is above target,
    nseq if nseq is below it)."""
    def f(neff):
        s = neff / nseq
        _, e = priorify(m, t_counts * s, e_counts * s, pri)
        return mean_match_relent(m, e) - etarget
    if f(min_neff) >= 0:
        return min_neff
    if f(nseq) <= 0:
        return float(nseq)
    lo, hi = 0.0, float(nseq)
    while hi - lo > 0.01:
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def target_relent(clen, nbps, esigma=45.0, re_target=0.59, re_target_nobp=0.38):
    """set_target_relent (src/cmbuild.c:2486-2500)."""
    re = re_target if nbps > 0 else re_target_nobp
    et = (esigma - math.log2(2.0 / (clen * (clen + 1.0)))) / clen
    return max(et, re)


def emit_map(m):
    """CreateEmitMap (src/display.c:1110): lpos/rpos per node in consensus
    coordinates."""
    lpos, rpos = [], []
    for nd_type, (i, j) in zip(m.nd_types, m.spans):
        lpos.append(i if nd_type in (MATP_nd, MATL_nd) else i - 1)
        rpos.append(j if nd_type in (MATP_nd, MATR_nd) else j + 1)
    return lpos, rpos


def detach_dual_inserts(m, t):
    """cm_find_and_detach_dual_inserts / cm_detach_state
    (src/cm_modelmaker.c:1800-2005): where an IL and an IR both insert after
    the same consensus column, the one followed by an END state is made
    unreachable (transitions into it from its node's split set set to 0 and
    the vectors renormalised)."""
    lpos, rpos = emit_map(m)
    lins, rins = {0: 1}, {m.clen: 2}
    for nd, typ in enumerate(m.nd_types):
        s = m.nodemap[nd]
        if typ == MATP_nd:
            lins[lpos[nd]] = s + 4; rins[rpos[nd] - 1] = s + 5
        elif typ == MATL_nd:
            lins[lpos[nd]] = s + 2
        elif typ == MATR_nd:
            rins[rpos[nd] - 1] = s + 2
        elif typ == BEGR_nd:
            lins[lpos[nd]] = s + 1
    for cc in range(m.clen + 1):
        if cc in lins and cc in rins:
            i1, i2 = lins[cc], rins[cc]
            to_detach = i1 if m.st[i1 + 1] == E_st else i2
            x_off = 0 if m.st[to_detach] == IL_st else 1
            for y in range(m.pnum[to_detach] - 1, 0, -1):
                x = m.plast[to_detach] - y
                t[x][x_off] = 0.0
                t[x, :m.cnum[x]] /= t[x, :m.cnum[x]].sum()
    return t


# ---------------------------------------------------------------------------
# writing and building

def _fmt(p, bg):
    return "*" if p <= 0 else "%.6f" % math.log2(p / bg)


def cm_text(m, t, e, name, W, bands, nseq, effn, el_self=0.94, pbegin=0.05, pend=0.05):
    """The model as an INFERNAL1/a text (the reference's CM file layout,
    src/cm_file.c:602-786) that infx_cm_ReadText reads back."""
    dmin1, dmax1, dmin2, dmax2 = bands
    lpos, rpos = emit_map(m)
    out = ["INFERNAL1/a [synthetic]", "NAME     %s" % name, "STATES   %d" % m.M, "NODES    %d" % len(m.nd_types),
           "CLEN     %d" % m.clen, "W        %d" % W, "ALPH     RNA", "RF       no", "CONS     yes", "MAP      yes",
           "PBEGIN   %g" % pbegin, "PEND     %g" % pend, "WBETA    1e-07", "QDBBETA1 1e-07", "QDBBETA2 1e-15",
           "N2OMEGA  1.52588e-05", "N3OMEGA  1.52588e-05", "ELSELF   %.8f" % math.log2(el_self),
           "NSEQ     %d" % nseq, "EFFN     %f" % effn, "CKSUM    0", "NULL     0.000  0.000  0.000  0.000", "CM"]
    first_of = {s: nd for nd, s in enumerate(m.nodemap)}
    for v in range(m.M):
        if v in first_of:
            nd = first_of[v]
            typ = m.nd_types[nd]
            lp = str(lpos[nd]) if typ in (MATP_nd, MATL_nd) else "-"
            rp = str(rpos[nd]) if typ in (MATP_nd, MATR_nd) else "-"
            out.append("%45s[ %-4s %4d ] %6s %6s - - - -" % ("", NODE_NAMES[typ], nd, lp, rp))
        s = m.st[v]
        fields = [STATE_NAMES[s], str(v), str(m.plast[v]), str(m.pnum[v]), str(m.cfirst[v]), str(m.cnum[v]),
                  str(dmin2[v]), str(dmin1[v]), str(dmax1[v]), str(dmax2[v])]
        if s != B_st:
            fields += [_fmt(t[v][k], 1.0) for k in range(m.cnum[v])]
        if s in (ML_st, MR_st, IL_st, IR_st):
            fields += [_fmt(e[v][x], 0.25) for x in range(4)]
        elif s == MP_st:
            fields += [_fmt(e[v][x], 1 / 16) for x in range(16)]
        out.append("    " + " ".join(fields))
    out.append("//")
    return "\n".join(out) + "\n"


def estimate(ax, ss_cons, pri, symfrac=0.5):
    """cmbuild's default path from a digital alignment (codes above) and its
    consensus structure to probabilities: PB weights, weighted consensus
    columns, structure knocked out of non-consensus columns, guide tree and
    rebalanced states, per-sequence parse counts, flanking-insert counts
    zeroed, entropy weighting, Dirichlet-mixture posterior means, dual
    inserts detached, insert emissions flat.  Sequences must not contain
    missing data ('~'); local-end columns are not used.
    Returns (model, t, e, neff, etarget)."""
    ax = np.asarray(ax, dtype=np.int64)
    if (ax == MISSING).any():
        raise ValueError("truncated (~) sequences are not supported")
    nseq, alen = ax.shape
    if len(ss_cons) != alen:
        raise ValueError("structure has %d columns, alignment %d" % (len(ss_cons), alen))
    wts = pb_weights(ax, symfrac)
    mat = match_columns(ax, wts, symfrac)
    ct = ss_to_ct(ss_cons)
    for col in range(1, alen + 1):
        if not mat[col]:
            if ct[col]:
                ct[ct[col]] = 0
            ct[col] = 0
    c2a = [0] + [col for col in range(1, alen + 1) if mat[col]]
    a2c = {col: c for c, col in enumerate(c2a) if c}
    ct_cons = [0] + [a2c[ct[col]] if ct[col] else 0 for col in c2a[1:]]
    m = Model(ct_cons)
    gorder = preorder(guide_tree(ct, mat))
    if [nd.type for nd, _ in gorder] != m.nd_types:
        raise AssertionError("alignment guide tree differs from the consensus one")
    tc = np.zeros((m.M, 6))
    ec = np.zeros((m.M, 16))
    for r in range(nseq):
        count_parse(m, gorder, ax[r], wts[r], tc, ec)
    zero_flanking_insert_counts(m, tc)
    nbps = sum(1 for typ in m.nd_types if typ == MATP_nd)
    etarget = target_relent(m.clen, nbps)
    neff = entropy_weight(m, tc, ec, nseq, pri, etarget)
    scale = float(np.float32(neff / nseq))
    t, e = priorify(m, tc * scale, ec * scale, pri)
    t = detach_dual_inserts(m, t)
    for v in range(m.M):
        if m.st[v] in (IL_st, IR_st):
            e[v, :] = 0.0
            e[v, :4] = 0.25
    return m, t, e, neff, etarget


def _bands_and_W(text0, M):
    import ctypes as C
    from . import CM, _lib
    cm = CM.from_text(text0)
    W = C.c_int()
    Z = C.c_int()
    eb = C.create_string_buffer(128)
    _lib.infx_CalculateQueryDependentBands.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_double,
                                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
    st = _lib.infx_CalculateQueryDependentBands(C.cast(cm.ptr, C.c_void_p), eb, C.cast(cm.c.qdbinfo, C.c_void_p),
                                                1e-7, C.byref(W), C.byref(Z))
    if st != 0:
        raise RuntimeError(eb.value.decode())
    q = cm.c.qdbinfo.contents
    bands = ([q.dmin1[v] for v in range(M)], [q.dmax1[v] for v in range(M)],
             [q.dmin2[v] for v in range(M)], [q.dmax2[v] for v in range(M)])
    return W.value, bands


def build_cm_from_alignment(ax, ss_cons, name, prior_path, el_self=0.94):
    """A CM (INFERNAL1/a text) from a digital alignment and its consensus
    structure with cmbuild's default options, and a dict of summary numbers.
    Not covered: --hand (RF-defined consensus), truncated sequences,
    alternative weighting or prior options."""
    pri = read_prior(prior_path)
    m, t, e, neff, etarget = estimate(ax, ss_cons, pri)
    nseq = np.asarray(ax).shape[0]
    zero = [0] * m.M
    full = [4 * m.clen] * m.M
    text0 = cm_text(m, t, e, name, 4 * m.clen, (zero, full, zero, full), nseq, neff, el_self)
    W, bands = _bands_and_W(text0, m.M)
    text = cm_text(m, t, e, name, W, bands, nseq, neff, el_self)
    info = {"clen": m.clen, "M": m.M, "nodes": len(m.nd_types), "W": W, "neff": neff, "etarget": etarget,
            "nbps": sum(1 for typ in m.nd_types if typ == MATP_nd),
            "nbif": sum(1 for typ in m.nd_types if typ == BIF_nd)}
    return text, info


def build_cm(ss, aln, name, prior_path, el_self=0.94):
    """A CM from a consensus structure and an ungapped alignment of canonical
    residues (codes 0..3, one column per consensus position)."""
    aln = np.asarray(aln, dtype=np.int64)
    if aln.shape[1] != len(ss):
        raise ValueError("alignment has %d columns, structure %d" % (aln.shape[1], len(ss)))
    return build_cm_from_alignment(aln, ss, name, prior_path, el_self)


# ---------------------------------------------------------------------------
# synthetic alignments

_PAIRS = [(0, 3), (3, 0), (1, 2), (2, 1), (2, 3), (3, 2)]    # AU UA CG GC GU UG
_PAIR_P = np.array([0.22, 0.22, 0.24, 0.24, 0.04, 0.04])


def stem_loop_structure(spec):
    """Dot-bracket string from a nested spec: ['.', n] an unpaired run,
    ['stem', bp, inner] a helix of bp pairs enclosing the spec list inner."""
    out = []
    for item in spec:
        if item[0] == ".":
            out.append("." * item[1])
        else:
            _, bp, inner = item
            out.append("<" * bp + stem_loop_structure(inner) + ">" * bp)
    return "".join(out)


def _loops(spec, acc):
    for item in spec:
        if item[0] == ".":
            acc.append(item)
        else:
            _loops(item[2], acc)
    return acc


def pad_to_length(spec, length, rng):
    """Grow randomly chosen unpaired runs of spec by one nucleotide at a time
    until the structure is <length> long."""
    ss = stem_loop_structure(spec)
    if len(ss) > length:
        raise ValueError("structure too long for %d nt: %d" % (length, len(ss)))
    loops = _loops(spec, [])
    for _ in range(length - len(ss)):
        loops[int(rng.integers(len(loops)))][1] += 1
    ss = stem_loop_structure(spec)
    assert len(ss) == length
    return ss


def cloverleaf_76():
    """76-nt tRNA-like cloverleaf: a 7-bp acceptor stem closing a four-way
    junction of a 4-bp D stem (8-nt loop), a 5-bp anticodon stem (7-nt loop)
    and a 5-bp T stem (7-nt loop), then a 4-nt 3' tail."""
    spec = [["stem", 7, [[".", 2], ["stem", 4, [[".", 8]]], [".", 1], ["stem", 5, [[".", 7]]], [".", 5],
                         ["stem", 5, [[".", 7]]]]], [".", 4]]
    ss = stem_loop_structure(spec)
    assert len(ss) == 76, len(ss)
    return ss


def multiloop_structure(length, nstems, rng, bp_range=(4, 10), junction=3, hairpin=(4, 8), internal=(2, 6)):
    """A closing stem around a <junction>-way multiloop whose branches are
    chains of stems joined by internal loops and capped by hairpins; nstems
    stems in total, loops grown at random to reach <length>."""
    per = [0] * junction
    for k in range(nstems - 1):
        per[k % junction] += 1

    def bp():
        return int(rng.integers(bp_range[0], bp_range[1] + 1))

    def branch(n):
        if n == 1:
            return [["stem", bp(), [[".", int(rng.integers(hairpin[0], hairpin[1] + 1))]]]]
        return [["stem", bp(), [[".", int(rng.integers(internal[0], internal[1] + 1))]] + branch(n - 1) +
                 [[".", int(rng.integers(0, internal[1] + 1))]]]]

    inner = []
    for k in range(junction):
        inner.append([".", int(rng.integers(2, 6))])
        inner += branch(per[k])
    inner.append([".", int(rng.integers(2, 6))])
    spec = [[".", 2], ["stem", bp(), inner], [".", 2]]
    return pad_to_length(spec, length, rng)


def nested_structure(length, nstems, rng, bp_range=(3, 12), hairpin=(4, 8)):
    """nstems stems of bp_range pairs placed as a random tree (each new stem
    goes inside an existing stem or beside the others at the top level),
    loops grown at random to reach <length>."""
    top = []
    bodies = [top]
    for _ in range(nstems):
        parent = bodies[int(rng.integers(len(bodies)))]
        body = []
        parent.append(["stem", int(rng.integers(bp_range[0], bp_range[1] + 1)), body])
        bodies.append(body)

    def render(items, toplevel):
        out = []
        for it in items:
            out.append([".", int(rng.integers(1, 4))])
            inner = render(it[2], False)
            if not it[2]:
                inner = [[".", int(rng.integers(hairpin[0], hairpin[1] + 1))]]
            out.append(["stem", it[1], inner])
        out.append([".", int(rng.integers(1, 4))])
        return out

    return pad_to_length(render(top, True), length, rng)


def synthetic_alignment(ss, nseq, mut, rng):
    """A consensus sequence on the structure (canonical pairs at paired
    columns) and nseq sequences, each mutated per column with probability
    <mut>; a mutated paired column changes its partner too, to another
    canonical pair (compensatory change)."""
    ct = ss_to_ct(ss)
    L = len(ss)
    cons = np.zeros(L, dtype=np.int64)
    for p in range(1, L + 1):
        q = ct[p]
        if q == 0:
            cons[p - 1] = rng.integers(0, 4)
        elif q > p:
            a, b = _PAIRS[rng.choice(len(_PAIRS), p=_PAIR_P)]
            cons[p - 1], cons[q - 1] = a, b
    aln = np.tile(cons, (nseq, 1))
    for s in range(nseq):
        for p in range(1, L + 1):
            q = ct[p]
            if rng.random() >= mut:
                continue
            if q == 0:
                aln[s, p - 1] = (aln[s, p - 1] + rng.integers(1, 4)) % 4
            else:
                cur = (aln[s, p - 1], aln[s, q - 1]) if p < q else (aln[s, q - 1], aln[s, p - 1])
                choices = [pr for pr in _PAIRS if pr != tuple(int(x) for x in cur)]
                a, b = choices[rng.integers(len(choices))]
                lo, hi = (p, q) if p < q else (q, p)
                aln[s, lo - 1], aln[s, hi - 1] = a, b
    return aln
