"""Host-side check of the truncated scan kernel's address arithmetic.

The profile RefTrCYKScan / RefITrInsideScan upload (state descriptors, row,
deck and T-row sizes, emission table, truncated-begin list and its per-block
lists) is taken from infx_debug_tr_profile, and the index formulas of
infernal_amd/csrc/scan_trunc.hip are evaluated for every state over its whole
band (the kernel clips a state's band to min(jp, W of the window, dx), so the
full band bounds every row and window length):

  rows      J, L and R rows share one layout: cell d of state c at
            base[c] + d, in [0, R)
  T rows    cell d of bifurcation c at tbase[c] + d, in [0, RT), inside
            c's own stretch [tbase[c], tbase[c] + dx[c]]
  decks     BEGL state c keeps a ring of ring_mask + 1 rows of ring_stride
            cells; cell d of ring row r at base[c] + r * stride + d, in
            [0, RB) for every r; a bifurcation reads rows j - k for k up to
            its BEGR band (kdx_y), so the ring must hold kdx_y + 1 rows, and
            the rings of different states must not overlap
  emission  [0, nesc) for residue codes 0..15

Every read of a child's cell (c, d') must also fall inside the band the
profile stored for c (lo[c] <= d' <= hi[c]); outside it the index would be
in bounds but name another state's cell.  The root's per-block begin lists
must name exactly the begins whose band reaches the block, in ascending
order.

This runs without a GPU, before any kernel change goes to one.
"""
import ctypes as C
import os

import numpy as np
import pytest

import infernal_amd as ia

CMDIR = os.path.join(os.path.dirname(__file__), "golden", "cm")
KP = 18
KIND_GENERIC, KIND_IL, KIND_B, KIND_BEGL, KIND_SKIP = 0, 1, 2, 3, 4
EMITLEFT, EMITRIGHT, EMITPAIR, EMITNONE = 0, 1, 2, 3

TSTATE = np.dtype([("kind", "<i4"), ("emitmode", "<i4"), ("cnum", "<i4"), ("dx", "<i4"), ("base", "<i4"),
                   ("tbase", "<i4"), ("esc_off", "<i4"), ("lesc_off", "<i4"), ("resc_off", "<i4"),
                   ("detached", "<i4"), ("first_r", "<i4"), ("init_ok", "<i4"), ("banded", "<i4"),
                   ("kdx_y", "<i4"), ("kdx_w", "<i4"), ("endsc", "<u4"), ("tsc", "<u4", 6),
                   ("child_base", "<i4", 6), ("ring_mask", "<i4"), ("ring_stride", "<i4"), ("level", "<i4"),
                   ("pad", "<i4")])
TRPEN = np.dtype([("y", "<i4"), ("base", "<i4"), ("tbase", "<i4"), ("dx", "<i4"), ("flags", "<i4"), ("pen", "<u4")])
assert TSTATE.itemsize == 128 and TRPEN.itemsize == 24


def tr_profile(cm, qdbidx, semi, pass_idx):
    lib = ia.lib()
    f = lib.infx_debug_tr_profile
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                  C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_char_p]
    M = cm.M
    st = np.zeros(M + 4, dtype=TSTATE)
    pen = np.zeros(M + 4, dtype=TRPEN)
    pcap = (M + 4) * ((cm.W + 63) // 64 + 2) + 64
    pblk = np.zeros(pcap, dtype=np.int32)
    lohi = np.zeros(2 * M, dtype=np.int32)
    dims = np.zeros(8, dtype=np.int32)
    eb = C.create_string_buffer(128)
    n = f(C.cast(cm.ptr, C.c_void_p), cm.c.trsmx, qdbidx, semi, pass_idx, st.ctypes.data, len(st),
          pen.ctypes.data, len(pen), pblk.ctypes.data, pcap, lohi.ctypes.data, dims.ctypes.data, eb)
    assert n >= 0, (n, eb.value)
    assert n == M and dims[5] <= len(pen) and dims[6] <= pcap
    return {"st": st[:n], "pen": pen[:dims[5]], "pblk": pblk[:dims[6]], "lo": lohi[0::2], "hi": lohi[1::2],
            "R": int(dims[0]), "RB": int(dims[1]), "RT": int(dims[2]), "nesc": int(dims[3]), "root_dx": int(dims[4])}


class Checker:
    """Collects out-of-range accesses; each access names the buffer, the
    owning state (or None) and the cell indices."""

    def __init__(self, p):
        self.p = p
        self.errors = []

    def _err(self, msg):
        if len(self.errors) < 20:
            self.errors.append(msg)

    def cells(self, buf, c, d, ctx):
        """Cells d (array) of state c in the J/L/R rows ('row') or the decks ('deck')."""
        d = np.atleast_1d(np.asarray(d))
        if d.size == 0:
            return
        p = self.p
        s = p["st"][c]
        base = int(s["base"])
        size = p["R"] if buf == "row" else p["RB"]
        if buf == "row":
            idx = base + d
        else:                                  # every ring row: first and last
            if int(s["ring_stride"]) != p["hi"][c] - p["lo"][c] + 1:
                self._err("deck stride %d of state %d is not its band width" % (int(s["ring_stride"]), c))
            last = int(s["ring_mask"]) * int(s["ring_stride"])
            idx = np.concatenate([base + d, base + last + d])
        if idx.min() < 0 or idx.max() >= size:
            self._err("%s index [%d, %d] outside [0, %d) for state %d %s" % (buf, idx.min(), idx.max(), size, c, ctx))
        if d.min() < p["lo"][c] or d.max() > p["hi"][c]:
            self._err("%s cells d [%d, %d] of state %d outside its band [%d, %d] %s"
                      % (buf, d.min(), d.max(), c, p["lo"][c], p["hi"][c], ctx))

    def tcells(self, c, d, ctx):
        d = np.atleast_1d(np.asarray(d))
        if d.size == 0:
            return
        s = self.p["st"][c]
        tb, dx = int(s["tbase"]), int(s["dx"])
        if d.min() < 0 or d.max() > dx or tb + d.max() >= self.p["RT"] or tb < 0:
            self._err("T cells [%d, %d] of state %d outside [0, %d] (tbase %d, RT %d) %s"
                      % (d.min(), d.max(), c, dx, tb, self.p["RT"], ctx))

    def esc(self, idx, ctx):
        idx = np.atleast_1d(np.asarray(idx))
        if idx.min() < 0 or idx.max() >= self.p["nesc"]:
            self._err("emission index [%d, %d] outside [0, %d) %s" % (idx.min(), idx.max(), self.p["nesc"], ctx))


def replay(cm, p, states=True):
    """Every access tr_scan_kernel makes for the states (when `states`) and
    the root, at the widest band (d = 1..dx).  The states' part does not
    depend on the truncation pass (only the begin list does)."""
    c = cm.c
    chk = Checker(p)
    codes = np.arange(16)
    for v in range(c.M - 1, 0, -1) if states else ():
        s = p["st"][v]
        kind, em, dx = int(s["kind"]), int(s["emitmode"]), int(s["dx"])
        if kind == KIND_SKIP or dx < 1:
            continue
        ctx = "(v %d kind %d em %d)" % (v, kind, em)
        ds = np.arange(1, dx + 1)
        kids = [c.cfirst[v] + k for k in range(int(s["cnum"]))]
        for k, ch in enumerate(kids):
            assert int(s["child_base"][k]) == int(p["st"][ch]["base"]), ctx
        if kind == KIND_B:
            w, y = c.cfirst[v], c.cnum[v]
            assert int(s["child_base"][0]) == int(p["st"][w]["base"]) and int(s["child_base"][1]) == int(p["st"][y]["base"])
            assert int(s["child_base"][2]) == int(p["st"][w]["ring_mask"]) and int(s["child_base"][3]) == int(p["st"][w]["ring_stride"])
            ring = int(p["st"][w]["ring_mask"]) + 1
            kdy = int(s["kdx_y"]) if s["banded"] else cm.W
            if ring < kdy + 1 or ring & (ring - 1):
                chk._err("deck ring of %d has %d rows for k up to %d %s" % (w, ring, kdy, ctx))
            for d in ds:
                if s["banded"]:
                    kmin, kmax = max(0, d - int(s["kdx_w"])), min(int(s["kdx_y"]), d)
                else:
                    kmin, kmax = 0, d
                if kmin <= kmax:
                    ks = np.arange(kmin, kmax + 1)
                    chk.cells("deck", w, d - ks, ctx)          # J and R decks of w at d - k
                    chk.cells("row", y, ks, ctx)               # J and L rows of y at k
                kn, kx = max(1, kmin), min(d - 1, kmax)
                if kn <= kx:                                   # T: R deck of w, L row of y
                    ks = np.arange(kn, kx + 1)
                    chk.cells("deck", w, d - ks, ctx)
                    chk.cells("row", y, ks, ctx)
                if kmin == 0:
                    chk.cells("deck", w, d, ctx)               # L: w alone
                if kmax == d:
                    chk.cells("row", y, d, ctx)                # R: y alone
            # a block walks k over the union of its lanes' ranges: every lane
            # loads the BEGR cell of each such k; lanes outside their range
            # read deck cell 0
            for d0 in range(1, dx + 1, 64):
                dl = min(d0 + 63, dx)
                kb0 = max(0, d0 - int(s["kdx_w"])) if s["banded"] else 0
                kb1 = min(int(s["kdx_y"]), dl) if s["banded"] else dl
                if kb0 <= kb1:
                    chk.cells("row", y, np.arange(kb0, kb1 + 1), ctx + " union")
                chk.cells("deck", w, 0, ctx + " idle")
            chk.cells("row", v, ds, ctx)
            chk.tcells(v, ds, ctx)
        elif kind == KIND_BEGL:
            for ch in kids:
                chk.cells("row", ch, ds, ctx)
            chk.cells("deck", v, ds, ctx)
        elif em == EMITLEFT:
            if s["detached"]:
                continue
            if kind == KIND_IL:
                chk.cells("row", v, 0, ctx)                    # the chain's carried d - 1 = 0 cell
            chk.cells("row", v, ds, ctx)
            for ch in kids:
                chk.cells("row", ch, ds - 1, ctx)              # J, L at d - 1 (idle lanes repeat d = dx)
            for ch in kids[int(s["first_r"]):]:
                chk.cells("row", ch, ds, ctx)                  # R mode at d
            chk.esc(int(s["esc_off"]) + codes, ctx)
        elif em == EMITRIGHT:
            if s["detached"]:
                continue
            chk.cells("row", v, ds, ctx)
            for ch in kids:
                chk.cells("row", ch, ds - 1, ctx)              # previous row, J and R
            for ch in kids[int(s["first_r"]):]:
                chk.cells("row", ch, ds, ctx)                  # L mode at d
            chk.esc(int(s["esc_off"]) + codes, ctx)
        elif em == EMITPAIR:
            chk.cells("row", v, ds, ctx)
            for ch in kids:
                if dx >= 2:
                    chk.cells("row", ch, ds[1:] - 2, ctx)      # J at d - 2, previous row
                chk.cells("row", ch, ds - 1, ctx)              # L (this row), R (previous row) at d - 1
            xi, xj = np.meshgrid(codes, codes)
            chk.esc(int(s["esc_off"]) + xi * KP + xj, ctx)
            chk.esc(int(s["lesc_off"]) + codes, ctx)
            chk.esc(int(s["resc_off"]) + codes, ctx)
        else:                                                  # S, D
            chk.cells("row", v, ds, ctx)
            for ch in kids:
                chk.cells("row", ch, ds, ctx)
    # root: per-block begin lists
    pen, pblk = p["pen"], p["pblk"]
    nblk = (cm.W + 63) // 64 + 1
    nb_used = (p["root_dx"] + 63) // 64
    assert nb_used <= nblk and len(pblk) >= nblk + 1
    assert pblk[0] == 0 and all(pblk[b] <= pblk[b + 1] for b in range(nblk))
    for b in range(nb_used):
        lst = [int(pblk[nblk + 1 + t]) for t in range(pblk[b], pblk[b + 1])]
        want = [k for k in range(len(pen)) if pen[k]["dx"] >= 64 * b + 1]
        if lst != want:
            chk._err("root block %d lists begins %s, want %s" % (b, lst[:8], want[:8]))
        for k in lst:
            q = pen[k]
            y = int(q["y"])
            assert int(q["base"]) == int(p["st"][y]["base"]) and int(q["dx"]) == int(p["st"][y]["dx"])
            hi = min(p["root_dx"], int(q["dx"]), 64 * b + 64)
            ds = np.arange(64 * b + 1, hi + 1)
            if ds.size:
                chk.cells("row", y, ds, "(root begin %d)" % y)
                chk.cells("row", y, 1, "(root begin %d, lanes past its band)" % y)
                if int(q["flags"]) & 1:
                    assert int(q["tbase"]) == int(p["st"][y]["tbase"])
                    chk.tcells(y, ds, "(root begin %d)" % y)
    # deck rings of different BEGL states do not overlap
    rings = sorted((int(p["st"][v]["base"]) + int(p["lo"][v]),
                    int(p["st"][v]["base"]) + int(p["st"][v]["ring_mask"]) * int(p["st"][v]["ring_stride"]) + int(p["hi"][v]))
                   for v in range(1, c.M) if p["st"][v]["kind"] == KIND_BEGL and p["lo"][v] <= p["hi"][v])
    for (a0, a1), (b0, _) in zip(rings, rings[1:]):
        if a1 >= b0:
            chk._err("deck rings overlap: [%d, %d] and %d" % (a0, a1, b0))
    # T stretches of different bifurcations do not overlap
    tb = sorted((int(p["st"][v]["tbase"]), int(p["st"][v]["dx"])) for v in range(1, c.M) if p["st"][v]["kind"] == KIND_B)
    for (t0, d0), (t1, _) in zip(tb, tb[1:]):
        if t0 + d0 >= t1:
            chk._err("T stretches overlap: %d + %d >= %d" % (t0, d0, t1))
    return chk


MODELS = ["tRNA.c.cm", "Cobalamin.c.cm", "synth_multiloop350.cm", "synth_nested1500.cm", "rnaseP-eubact.c.cm"]
PASSES = [ia.PLI_PASS_5P_AND_3P_FORCE, ia.PLI_PASS_5P_ONLY_FORCE, ia.PLI_PASS_3P_ONLY_FORCE]


@pytest.mark.parametrize("name", MODELS)
def test_trunc_kernel_addresses_in_bounds(name):
    for local in (True, False):
        cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=local, trunc=True, scanmx=True, trscanmx=True)
        for semi, qdb in ((0, 1), (1, 2), (0, 0)):
            for k, pidx in enumerate(PASSES):
                p = tr_profile(cm, qdb, semi, pidx)
                chk = replay(cm, p, states=(k == 0))
                assert not chk.errors, (name, local, semi, qdb, pidx, chk.errors)


def test_trunc_checker_catches_planted_overreach():
    """The replay must flag a band one cell too narrow, a begin left out of a
    block's list and an emission table that ends before a singlet emitter's
    code-15 entry."""
    cm = ia.CM.read(os.path.join(CMDIR, "Cobalamin.c.cm")).configure(local=True, trunc=True, scanmx=True, trscanmx=True)
    p = tr_profile(cm, 1, 0, ia.PLI_PASS_5P_AND_3P_FORCE)
    assert not replay(cm, p).errors
    c = cm.c
    b = next(v for v in range(1, c.M) if p["st"][v]["kind"] == KIND_B)
    q = dict(p, hi=p["hi"].copy())
    q["hi"][b] = p["st"][b]["dx"] - 1                         # a bifurcation's own top cell
    assert any("outside its band" in e for e in replay(cm, q).errors)
    q = dict(p, pblk=p["pblk"].copy())
    nblk = (cm.W + 63) // 64 + 1
    q["pblk"][1:nblk + 1] -= 1                                # block 0 loses its last begin
    assert any("root block 0" in e for e in replay(cm, q).errors)
    e = next(v for v in range(1, c.M) if p["st"][v]["emitmode"] in (EMITLEFT, EMITRIGHT) and not p["st"][v]["detached"]
             and p["st"][v]["kind"] != KIND_SKIP and p["st"][v]["dx"] >= 1)
    q = dict(p, nesc=int(p["st"][e]["esc_off"]) + 15)
    assert any("emission index" in e for e in replay(cm, q).errors)


def tr_schedule(cm, qdbidx, semi, pass_idx, nw, vsc):
    lib = ia.lib()
    f = lib.infx_debug_tr_schedule
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                  C.c_int, C.POINTER(C.c_int), C.c_char_p]
    cap = 4 * cm.M * ((cm.W + 63) // 64 + 1) + 64
    sched = np.zeros(cap, dtype=np.int32)
    items = np.zeros(3 * cap, dtype=np.int32)
    nlev = C.c_int()
    eb = C.create_string_buffer(128)
    n = f(C.cast(cm.ptr, C.c_void_p), cm.c.trsmx, qdbidx, semi, pass_idx, nw, int(vsc), sched.ctypes.data, cap,
          items.ctypes.data, len(items), C.byref(nlev), eb)
    assert n >= 0, (n, eb.value)
    L = nlev.value
    return L, sched[:L * nw + 1], items[:3 * n].reshape(n, 3)


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Cobalamin.c.cm", "synth_nested1500.cm"])
def test_trunc_row_schedule(name):
    """The row schedule the truncated kernel runs: every state with cells is
    covered, block by block, exactly once; an item sits at its state's level;
    every child whose same-row cells a state reads (B: BEGL and BEGR) is at a
    lower level, so the barrier after each level orders every same-row read
    after its write; IL chains (and every state, when per-state maxima are
    wanted) are one item on one wavefront."""
    cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=True, trunc=True, scanmx=True, trscanmx=True)
    c = cm.c
    p = tr_profile(cm, 1, 0, ia.PLI_PASS_5P_AND_3P_FORCE)
    st = p["st"]
    lev = st["level"]
    for v in range(1, c.M):
        if lev[v] < 0:
            assert st[v]["kind"] == KIND_SKIP or st[v]["dx"] < 1
            continue
        kids = [c.cfirst[v], c.cnum[v]] if st[v]["kind"] == KIND_B else [c.cfirst[v] + k for k in range(c.cnum[v])]
        for ch in kids:
            if ch != v and lev[ch] >= 0:
                assert lev[ch] < lev[v], (v, ch, lev[v], lev[ch])
    for nw, vsc in ((1, False), (4, False), (8, False), (4, True)):
        L, sched, items = tr_schedule(cm, 1, 0, ia.PLI_PASS_5P_AND_3P_FORCE, nw, vsc)
        assert L == lev.max() + 1 and sched[0] == 0 and sched[-1] == len(items)
        assert all(sched[i] <= sched[i + 1] for i in range(len(sched) - 1))
        cover = {}
        for l in range(L):
            for wv in range(nw):
                for v, b0, b1 in items[sched[l * nw + wv]:sched[l * nw + wv + 1]]:
                    assert lev[v] == l and 0 <= b0 < b1
                    if vsc or st[v]["kind"] == KIND_IL:
                        assert (b0, b1) == (0, (st[v]["dx"] + 63) // 64), (v, b0, b1)
                    for b in range(b0, b1):
                        assert (v, b) not in cover
                        cover[(v, b)] = True
        want = {(v, b) for v in range(1, c.M) if lev[v] >= 0 for b in range((st[v]["dx"] + 63) // 64)}
        assert set(cover) == want
