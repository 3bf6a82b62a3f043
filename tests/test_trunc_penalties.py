"""Truncated-begin penalties (cm->trp, built by cm_Configure with truncation;
a restatement of cm_tr_penalties_Create, src/cm_trunc.c:51-450) pinned by
what they mean, not by a second copy of the loop.  Both the GPU scanners and
the oracle read this table, so GPU-vs-oracle parity cannot catch a mistake
in it; these checks can.

The penalty of a state is log2 of the probability that a truncated
alignment is rooted there.  For each node of type MATP, MATL, MATR or BIF
with first state m, that probability p(node) is shared between m and the
insert states of the node before it (InsertsGivenNodeIndex,
src/cm.c:4897) in proportion to their expected occupancy psi.  So:

* global, 5' and 3' truncation: every group sums to 2 / (clen (clen+1)),
  the probability of one of the clen (clen+1) / 2 equiprobable fragments;
  5' only (3' only): 1 / clen for nodes whose consensus span reaches clen
  (starts at 1), no penalty otherwise (src/cm_trunc.c:110-150, :320-340);
* local: the reference's own check, cm_tr_penalties_Validate
  (src/cm_trunc.c:475-628), on the table built with its ignore_inserts flag
  (the flag the check assumes; infx_debug_tr_penalties): over every fragment
  g..h that the truncation mode can make, the probability of the lowest node
  spanning it sums to 1 (to 1e-4, the reference's tolerance).  The table
  cm_Configure builds is the same computation with the inserts' shares
  taken out of each node's probability;
* the shares: exp2(penalty) of an insert over that of its match state equals
  psi(insert) / psi(match), with psi solved here as a linear system over the
  global model's transitions (a bifurcation enters both children), not by
  the library's recursion;
* the integer penalties are Prob2Score(p, 1) = floor(0.5 + 1000 log2 p).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import infernal_amd as ia
from helpers import CMDIR

MODELS = ["tRNA.c.cm", "Vault.c.cm", "5S_rRNA.c.cm", "Cobalamin.c.cm", "Plant_SRP.c.cm", "rnaseP-eubact.c.cm"]
BIF_nd, MATP_nd, MATL_nd, MATR_nd, BEGL_nd, BEGR_nd, ROOT_nd, END_nd = range(8)
T53, T5, T3 = 0, 1, 2          # TRPENALTY_5P_AND_3P, TRPENALTY_5P_ONLY, TRPENALTY_3P_ONLY
S_st, E_st, B_st = 6, 7, 8      # state types (src/infernal.h)
IMP = -1e36


class TRP(C.Structure):
    _fields_ = [("M", C.c_int), ("ignored_inserts", C.c_int),
                ("g_ptyAA", C.POINTER(C.POINTER(C.c_float))), ("l_ptyAA", C.POINTER(C.POINTER(C.c_float))),
                ("ig_ptyAA", C.POINTER(C.POINTER(C.c_int))), ("il_ptyAA", C.POINTER(C.POINTER(C.c_int)))]


def inserts_of_node(nodemap, ndtype, nd):
    """src/cm.c:4897 InsertsGivenNodeIndex: the insert states a node owns."""
    if nd < 0:
        return []
    v = nodemap[nd]
    t = ndtype[nd]
    if t == MATP_nd:
        return [v + 4, v + 5]
    if t in (MATL_nd, MATR_nd, BEGR_nd):
        return [v + 1 + (t != BEGR_nd)]
    if t == ROOT_nd:
        return [v + 1, v + 2]
    return []


def occupancy(c):
    """Expected number of visits to each state in the global model: psi =
    e_0 + T' psi, with T the transition probabilities (a bifurcation enters
    its BEGL and BEGR children with probability 1)."""
    M = c.M
    T = np.zeros((M, M))
    for x in range(M):
        st = ord(c.sttype[x])
        if st == E_st:
            continue
        if st == B_st:
            T[x, c.cfirst[x]] = 1.0
            T[x, c.cnum[x]] = 1.0
            continue
        for k in range(c.cnum[x]):
            T[x, c.cfirst[x] + k] += c.t[x][k]
    e0 = np.zeros(M)
    e0[0] = 1.0
    return np.linalg.solve(np.eye(M) - T.T, e0)


def groups(c):
    """(node, consensus span lpos..rpos, [m, inserts...]) for every node a
    truncated alignment can be rooted at."""
    em = c.emap.contents
    ndtype = [ord(c.ndtype[n]) for n in range(c.nodes)]
    nodemap = [c.nodemap[n] for n in range(c.nodes)]
    out = []
    for nd in range(c.nodes):
        t = ndtype[nd]
        if t not in (MATP_nd, MATL_nd, MATR_nd, BIF_nd):
            continue
        lpos = em.lpos[nd] if t in (MATP_nd, MATL_nd) else em.lpos[nd] + 1
        rpos = em.rpos[nd] if t in (MATP_nd, MATR_nd) else em.rpos[nd] - 1
        out.append((nd, lpos, rpos, [nodemap[nd]] + inserts_of_node(nodemap, ndtype, nd - 1)))
    return out


def penalties(name):
    glob = ia.CM.read(os.path.join(CMDIR, name))                 # the file's (global) probabilities
    cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=True, trunc=True)
    trp = C.cast(cm.c.trp, C.POINTER(TRP)).contents
    M = cm.c.M
    tab = {}
    for key, arr in (("g", trp.g_ptyAA), ("l", trp.l_ptyAA), ("ig", trp.ig_ptyAA), ("il", trp.il_ptyAA)):
        tab[key] = np.array([[arr[i][v] for v in range(M)] for i in range(3)], dtype=np.float64)
    return glob, cm, tab


def group_prob(pty, idx, members):
    return sum(2.0 ** pty[idx][s] for s in members if pty[idx][s] > -1e30)


@pytest.mark.parametrize("name", MODELS)
def test_global_penalties_closed_forms(name):
    glob, cm, tab = penalties(name)
    clen = cm.c.clen
    g = tab["g"]
    n = 0
    for nd, lpos, rpos, mem in groups(cm.c):
        assert math.isclose(group_prob(g, T53, mem), 2.0 / (clen * (clen + 1)), rel_tol=2e-6), (name, nd)
        if rpos == clen:
            assert math.isclose(group_prob(g, T5, mem), 1.0 / clen, rel_tol=2e-6), (name, nd)
        else:
            assert all(g[T5][s] < -1e30 for s in mem), (name, nd)
        if lpos == 1:
            assert math.isclose(group_prob(g, T3, mem), 1.0 / clen, rel_tol=2e-6), (name, nd)
        else:
            assert all(g[T3][s] < -1e30 for s in mem), (name, nd)
        n += 1
    assert n > 10


def debug_tables(glob, ignore_inserts):
    """infx_debug_tr_penalties on an unconfigured model: (glocal, local)
    float tables, each [3][M]."""
    lib = ia.lib()
    f = lib.infx_debug_tr_penalties
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_char_p]
    M = glob.c.M
    out = np.zeros((2, 3, M), dtype=np.float32)
    eb = C.create_string_buffer(128)
    assert f(C.cast(glob.ptr, C.c_void_p), int(ignore_inserts), out.ctypes.data, eb) == 0, eb.value
    return out[0].astype(np.float64), out[1].astype(np.float64)


@pytest.mark.parametrize("name", MODELS)
def test_configured_table_is_the_debug_build(name):
    """The table cm_Configure attaches equals infx_debug_tr_penalties with
    the inserts' shares (so the checks on the debug build below cover the
    configured code path)."""
    glob, cm, tab = penalties(name)
    g, l = debug_tables(glob, 0)
    assert np.array_equal(g.astype(np.float32), tab["g"].astype(np.float32))
    assert np.array_equal(l.astype(np.float32), tab["l"].astype(np.float32))


@pytest.mark.parametrize("name", MODELS)
def test_local_penalties_sum_to_one_over_fragments(name):
    """cm_tr_penalties_Validate (src/cm_trunc.c:498-628), with its
    ignore_inserts flag: each fragment g..h is rooted at the lowest
    MATP/MATL/MATR/BIF node whose consensus span holds it (5' and 3'), ends
    at h (5' only) or starts at g (3' only); summed over the fragments, the
    probability of the root node's state is 1."""
    glob, cm, tab = penalties(name)
    clen = cm.c.clen
    _, l = debug_tables(glob, 1)
    gr = groups(cm.c)                      # ascending node index
    for idx, cond in ((T53, "both"), (T5, "5p"), (T3, "3p")):
        root = np.full((clen + 2, clen + 2), -1, dtype=np.int64)   # root[g][h]: group index
        for k, (nd, lpos, rpos, mem) in enumerate(gr):
            # later (lower) nodes overwrite: the lowest spanning node wins
            if cond == "both":
                root[lpos:, :rpos + 1] = k
            elif cond == "5p":
                root[lpos:, rpos] = k
            else:
                root[lpos, :rpos + 1] = k
        gg, hh = np.triu_indices(clen + 1)
        sel = (gg >= 1) & (hh >= gg)
        ks = root[gg[sel], hh[sel]]
        if cond == "both":
            assert (ks >= 0).all(), (name, "a fragment with no spanning node")
        ks = ks[ks >= 0]
        probs = np.array([2.0 ** l[idx][mem[0]] for (_, _, _, mem) in gr])
        total = probs[ks].sum()
        assert abs(total - 1.0) < 1e-4, (name, cond, total)


def test_validator_catches_a_wrong_penalty():
    """The fragment sum is sensitive: halving one node's probability in the
    table moves it past the tolerance."""
    glob, cm, tab = penalties("tRNA.c.cm")
    clen = cm.c.clen
    _, l = debug_tables(glob, 1)
    gr = groups(cm.c)
    root = np.full((clen + 2, clen + 2), -1, dtype=np.int64)
    for k, (nd, lpos, rpos, mem) in enumerate(gr):
        root[lpos:, :rpos + 1] = k
    gg, hh = np.triu_indices(clen + 1)
    sel = (gg >= 1) & (hh >= gg)
    ks = root[gg[sel], hh[sel]]
    probs = np.array([2.0 ** l[T53][mem[0]] for (_, _, _, mem) in gr])
    assert abs(probs[ks].sum() - 1.0) < 1e-4
    counts = np.bincount(ks, minlength=len(gr))
    k = int(np.argmax(counts * probs))        # the node that carries the most probability
    probs[k] *= 0.5
    assert abs(probs[ks].sum() - 1.0) > 1e-3


@pytest.mark.parametrize("name", MODELS)
def test_insert_shares_follow_occupancy(name):
    glob, cm, tab = penalties(name)
    psi = occupancy(glob.c)
    # the linear system's own sanity: every S state and every bifurcation is entered once
    st = glob.sttype()
    for v in range(glob.c.M):
        if st[v] in (S_st, B_st):
            assert abs(psi[v] - 1.0) < 1e-5, (name, v, psi[v])   # the file's rounded transitions leak ~1e-7
    checked = 0
    for nd, lpos, rpos, mem in groups(cm.c):
        m = mem[0]
        for i in mem[1:]:
            for key in ("g", "l"):
                for idx in (T53, T5, T3):
                    a, b = tab[key][idx][i], tab[key][idx][m]
                    if a < -1e30 or b < -1e30:
                        continue
                    want = math.log2(psi[i] / psi[m])
                    assert abs((a - b) - want) < 2e-4, (name, nd, key, idx, a - b, want)
                    checked += 1
    assert checked > 20


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Cobalamin.c.cm"])
def test_integer_penalties(name):
    """ig/il = Prob2Score(p, 1.0) = floor(0.5 + 1000 log2 p) of the same p
    whose log2 the float table holds (to float rounding of that log)."""
    glob, cm, tab = penalties(name)
    for f, i in (("g", "ig"), ("l", "il")):
        fl, it = tab[f], tab[i]
        on = fl > -1e30
        assert (it[~on] <= -987654321).all()
        assert np.abs(it[on] - 1000.0 * fl[on]).max() <= 0.5 + 1e-3
