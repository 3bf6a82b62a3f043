"""Brute-force enumeration of truncated parse trees of tiny CMs.

A truncated alignment of x[i..j] enters the model through a truncated begin
into some state y (scored by the penalty table cm->trp) in one of four modes:

  J  the subtree of y is aligned in full (a standard parse; local ends allowed)
  L  x[i..j] is a prefix of what the subtree of y emits: the target ends
     inside the subtree and everything right of j is missing
  R  x[i..j] is a suffix: everything left of i is missing
  T  (bifurcations) the left subtree is in R mode and the right one in L mode

The derivation rules, per state type (x_i the first, x_j the last residue of
the span, "end" a local end into EL that emits the rest of the span):

  L mode
    ML, IL   emit x_i; a one-residue span stops there; longer spans continue
             with a local end or a child in L mode (IL may loop on itself)
    MR, IR   silent (their residue is missing): local end, or a child in J or
             L mode on the same span (an IR never loops on itself here)
    MP       emits x_i with the left-marginal score; one residue stops;
             longer spans continue with a local end or a child in J or L mode
    S, D     local end, or a child in L mode on the same span
    B        local end; or the left child in J mode and the right child in L
             mode on a split with a non-empty right part; or the left child
             alone (J or L mode) on the whole span
  R mode     the mirror image (MR/IR/MP emit x_j; ML/IL are silent; B: left
             child in R mode and right child in J mode, or the right child
             alone in J or R mode)
  T mode     B only: local end, or the left child in R mode and the right
             child in L mode, both non-empty
  Detached inserts (an IL/IR followed by END_E) take part in no alignment.

This is the truncated grammar the scanner recursions tabulate
(src/cm_dpsearch_trunc.c:276-443 for TrCYK, :900-1130 for TrInside); the
enumeration is checked against the reference's own hand-listed paths of
src/itest_brute.c (score_brute_matl_cm) and then used as the brute-force
answer for the oracle's truncated scanners.

The enumeration works over an "algebra": numeric (a derivation is its float
score) or symbolic (a derivation is the tuple of its parameter names, as in
src/itest_brute.c).  For the integer Inside the scanner combines a few
alternatives by max instead of by sum (src/cm_dpsearch_trunc.c:1086-1091 for
MP, and the B-state L/R extensions); `inside=True` applies those collapses.
It also initialises the empty-span J cells of S, D and BEGL states from the
float local-end score assigned to an integer cell (src/cm_mx.c:6941,
:6955): that alternative is worth trunc(endsc) thousandths of a bit there.
"""
import math
from collections import Counter

from tinycm import IMP, log2sumexp2

D_ST, MP_ST, ML_ST, MR_ST, IL_ST, IR_ST, S_ST, E_ST, B_ST = range(9)
END_E = 18
TRMODE_T, TRMODE_R, TRMODE_L, TRMODE_J = 0, 1, 2, 3
KP = 18


class TrTables:
    """Configured score tables of a truncation-configured CM_t."""

    def __init__(self, cm, pty_idx):
        import ctypes as C
        c = cm.c
        M = c.M
        self.M = M
        self.sttype = [ord(c.sttype[v]) for v in range(M)]
        self.stid = [ord(c.stid[v]) for v in range(M)]
        self.cfirst = [c.cfirst[v] for v in range(M)]
        self.cnum = [c.cnum[v] for v in range(M)]
        self.tsc = [[c.tsc[v][k] for k in range(6)] for v in range(M)]
        self.endsc = [c.endsc[v] for v in range(M)]
        self.el = c.el_selfsc
        self.local = bool(c.flags & (1 << 10))   # CMH_LOCAL_BEGIN
        self.esc, self.lmesc, self.rmesc = [], [], []
        for v in range(M):
            t = self.sttype[v]
            n = KP * KP if t == MP_ST else KP if t in (ML_ST, MR_ST, IL_ST, IR_ST) else 0
            self.esc.append([c.oesc[v][x] for x in range(n)] if n else None)
            self.lmesc.append([c.lmesc[v][x] for x in range(KP)] if t == MP_ST else None)
            self.rmesc.append([c.rmesc[v][x] for x in range(KP)] if t == MP_ST else None)

        class TRP(C.Structure):
            _fields_ = [("M", C.c_int), ("ignored_inserts", C.c_int),
                        ("g_ptyAA", C.POINTER(C.POINTER(C.c_float))), ("l_ptyAA", C.POINTER(C.POINTER(C.c_float))),
                        ("ig_ptyAA", C.POINTER(C.POINTER(C.c_int))), ("il_ptyAA", C.POINTER(C.POINTER(C.c_int)))]

        trp = C.cast(c.trp, C.POINTER(TRP)).contents
        row = (trp.l_ptyAA if self.local else trp.g_ptyAA)[pty_idx]
        self.pty = [row[v] for v in range(M)]

    def detached(self, v):
        return self.sttype[v] in (IL_ST, IR_ST) and v + 1 < self.M and self.stid[v + 1] == END_E

    def children(self, v):
        return range(self.cfirst[v], self.cfirst[v] + self.cnum[v])


class NumAlg:
    """Derivation = float score (bits)."""

    def __init__(self, tb, x):
        self.tb, self.x = tb, x

    one = 0.0

    @staticmethod
    def cat(a, b):
        return a + b

    def t(self, v, c):
        return self.tb.tsc[v][c - self.tb.cfirst[v]]

    def em(self, v, i):
        return self.tb.esc[v][self.x[i]]

    def ep(self, v, i, j):
        return self.tb.esc[v][self.x[i] * KP + self.x[j]]

    def lm(self, v, i):
        return self.tb.lmesc[v][self.x[i]]

    def rm(self, v, j):
        return self.tb.rmesc[v][self.x[j]]

    def end(self, v, n):
        return self.tb.endsc[v] + self.tb.el * n

    @staticmethod
    def ok(a):
        return a > IMP


class SymAlg:
    """Derivation = tuple of parameter names, in the naming of
    src/itest_brute.c: t<v>t<y> transitions, msc / isc match and insert
    emissions, end<v> a local end, and elsc * el_self per residue EL emits."""

    def __init__(self, tb):
        self.tb = tb

    one = ()

    @staticmethod
    def cat(a, b):
        return a + b

    def t(self, v, c):
        return ("t%dt%d" % (v, c),)

    def _e(self, v):
        return ("isc",) if self.tb.sttype[v] in (IL_ST, IR_ST) else ("msc",)

    def em(self, v, i):
        return self._e(v)

    def ep(self, v, i, j):
        return ("msc",)

    lm = ep
    rm = ep

    def end(self, v, n):
        return ("end%d" % v,) + ("elsc", "el_self") * n

    @staticmethod
    def ok(a):
        return True


class TruncEnum:
    def __init__(self, tb, alg, inside=False):
        self.tb, self.alg, self.inside = tb, alg, inside
        self.memo = {}

    # -- helpers -----------------------------------------------------------
    def _finite(self, v, attr):
        return getattr(self.tb, attr)[v] > IMP

    def _endok(self, v):
        return self.tb.endsc[v] > IMP

    def _best_group(self, groups):
        """Inside: keep the group (list of derivations) with the larger sum,
        the scanner's max over alternatives; otherwise their union."""
        groups = [g for g in groups if g]
        if not groups:
            return []
        if not self.inside:
            return [a for g in groups for a in g]
        return max(groups, key=log2sumexp2)

    def _prefix(self, pre, derivs):
        if not self.alg.ok(pre):
            return []
        cat = self.alg.cat
        return [cat(pre, a) for a in derivs]

    def _pairs(self, left, right):
        cat = self.alg.cat
        return [cat(a, b) for a in left for b in right]

    # -- enumeration ---------------------------------------------------------
    def gen(self, v, mode, i, j):
        key = (v, mode, i, j)
        if key not in self.memo:
            out = self._gen(v, mode, i, j)
            if self.inside and mode == "J" and j < i and len(out) > 1:
                out = [max(out)]   # d = 0 cells are initialised by max
            self.memo[key] = out
        return self.memo[key]

    def _gen(self, v, mode, i, j):
        tb, A = self.tb, self.alg
        st = tb.sttype[v]
        d = j - i + 1
        if tb.detached(v) or d < 0:
            return []
        if mode != "J" and d < 1:
            return []
        if st == E_ST:
            return [A.one] if (mode == "J" and d == 0) else []
        out = []
        if st == B_ST:
            w, y = tb.cfirst[v], tb.cnum[v]
            if self._endok(v):
                out.append(A.end(v, d))
            if mode == "J":
                for k in range(i - 1, j + 1):
                    out += self._pairs(self.gen(w, "J", i, k), self.gen(y, "J", k + 1, j))
            elif mode == "L":
                for k in range(i - 1, j):   # right part k+1..j non-empty
                    out += self._pairs(self.gen(w, "J", i, k), self.gen(y, "L", k + 1, j))
                out += self._best_group([self.gen(w, "J", i, j), self.gen(w, "L", i, j)])
            elif mode == "R":
                for k in range(i, j + 1):   # left part i..k non-empty
                    out += self._pairs(self.gen(w, "R", i, k), self.gen(y, "J", k + 1, j))
                out += self._best_group([self.gen(y, "J", i, j), self.gen(y, "R", i, j)])
            else:  # T
                for k in range(i, j):
                    out += self._pairs(self.gen(w, "R", i, k), self.gen(y, "L", k + 1, j))
            return out
        if mode == "T":
            return []
        kids = list(tb.children(v))
        if st in (ML_ST, IL_ST, MR_ST, IR_ST, MP_ST):
            left = st in (ML_ST, IL_ST)
            right = st in (MR_ST, IR_ST)
            if mode == "J":
                if st == MP_ST:
                    if d < 2:
                        return []
                    e, ii, jj = A.ep(v, i, j), i + 1, j - 1
                elif left:
                    if d < 1:
                        return []
                    e, ii, jj = A.em(v, i), i + 1, j
                else:
                    if d < 1:
                        return []
                    e, ii, jj = A.em(v, j), i, j - 1
                if self._endok(v):
                    out += self._prefix(e, [A.end(v, jj - ii + 1)])
                for c in kids:
                    out += self._prefix(A.cat(e, A.t(v, c)), self.gen(c, "J", ii, jj))
                return out
            emits = (mode == "L" and (left or st == MP_ST)) or (mode == "R" and (right or st == MP_ST))
            if emits:
                if st == MP_ST:
                    e = A.lm(v, i) if mode == "L" else A.rm(v, j)
                else:
                    e = A.em(v, i) if mode == "L" else A.em(v, j)
                if not A.ok(e):
                    return []
                if d == 1:
                    return [e]
                ii, jj = (i + 1, j) if mode == "L" else (i, j - 1)
                if self._endok(v):
                    out += self._prefix(e, [A.end(v, d - 1)])
                for c in kids:
                    pre = A.cat(e, A.t(v, c))
                    if st == MP_ST:
                        out += self._prefix(pre, self._best_group([self.gen(c, "J", ii, jj), self.gen(c, mode, ii, jj)]))
                    else:
                        out += self._prefix(pre, self.gen(c, mode, ii, jj))
                return out
            # silent in this mode: the state's residue is the missing one
            if self._endok(v):
                out.append(A.end(v, d))
            for c in kids:
                if c == v:
                    continue   # no self-loop of a silent insert
                pre = A.t(v, c)
                out += self._prefix(pre, self.gen(c, "J", i, j) + self.gen(c, mode, i, j))
            return out
        # S, D (BEGL, BEGR included)
        if v != 0 and self._endok(v):
            if self.inside and d == 0:
                out.append(math.trunc(tb.endsc[v]) / 1000.0)   # float endsc in an int cell
            else:
                out.append(A.end(v, d))
        for c in kids:
            out += self._prefix(A.t(v, c), self.gen(c, mode, i, j))
        return out


def mode_maxima(tb, x, i, j, modes, inside):
    """Per mode, the best truncated begin max_y agg(derivations) + penalty[y]
    over x[i..j], and the first y attaining it (agg = max or log-sum)."""
    en = TruncEnum(tb, NumAlg(tb, x), inside=inside)
    out = {}
    for m in modes:
        best, arg, vals = None, None, {}
        for y in range(1, tb.M):
            if tb.pty[y] <= IMP:
                continue
            if m == "T" and tb.sttype[y] != B_ST:
                continue
            ds = en.gen(y, m, i, j)
            ds = [a for a in ds if a > IMP]
            if not ds:
                continue
            val = (log2sumexp2(ds) if inside else max(ds)) + tb.pty[y]
            vals[y] = val
            if best is None or val > best:
                best, arg = val, y
        out[m] = (best, arg, vals)
    return out


def symbolic_paths(tb, L, modes=("J", "L", "R")):
    """All truncated-begin derivations of an L-residue sequence (a single
    residue type), as a Counter of sorted parameter-name tuples per mode;
    truncated begins into every non-detached emitting state, at no cost."""
    en = TruncEnum(tb, SymAlg(tb))
    out = {}
    for m in modes:
        cnt = Counter()
        for y in range(1, tb.M):
            if tb.sttype[y] not in (MP_ST, ML_ST, MR_ST, IL_ST, IR_ST) or tb.detached(y):
                continue
            for p in en.gen(y, m, 1, L):
                cnt[tuple(sorted(p))] += 1
        out[m] = cnt
    return out
