"""Host-side check of the standard scan kernel's address arithmetic.

The scan profile the library uploads (state descriptors, row/deck sizes,
emission table, root band) is taken from infx_debug_std_profile, and the
index formulas of infernal_amd/csrc/scan_std.hip (multi-row sweeps, 64-lane
blocks, idle-lane stand-in cells, the bifurcation k-loop with its 4-wide
overshoot, local-begin accumulators, root, emission rows, residue reads) are
evaluated for every state over the chunk's warm-up and steady rows and at the
target's end.  Every access must fall inside its buffer:

  residues  device bytes [61, L + 128): the target itself, the 64 zero
            bytes after it, and up to NR-1 = 3 of the 64 zero bytes before
            it (lanes past a visit's widest band read at its edge, at most
            NR-1 residues before a chunk's first row)
  rows      [0, R) of each ring row (the kernel adds the ring row's offset):
            lanes past the band read the band's edge cell (Inside) or the
            stand-in cell dn (CYK)
  decks     [-G, RB + G) with G = W + 128 rounded up to 64 (guard cells)
  LDS       accumulator index [0, acc_n)
  emission  [0, nesc)

The replay walks the 64-cell blocks of a state one at a time.  The kernels
that take two or four blocks per step (Inside with the LDS table, the
wide-band CYK kernel) visit the same blocks with the same per-block masks,
which test_multi_block_steps_visit_the_same_blocks checks; the kernels with
the local-begin maxima in the global slot use the same accumulator indices
(row j+t at [t acc_n, (t+1) acc_n) of the slot's region).  A visit of NR
rows accesses, for each of its rows, what the two-row replay below checks for
a row at the same chunk position (the rows of one visit differ only in jp),
plus the residue j0+t of each idle row past the chunk's end.

This runs without a GPU, before any kernel change goes to one.
"""
import ctypes as C
import os

import numpy as np
import pytest

import infernal_amd as ia

CMDIR = os.path.join(os.path.dirname(__file__), "golden", "cm")
KP = 18
NR_MAX = 4          # most rows a kernel computes per state visit (scan_kernels.h INFX_NR_MAX)
KIND_GENERIC, KIND_IL, KIND_B, KIND_BEGL = 0, 1, 2, 3
EMITLEFT, EMITRIGHT, EMITPAIR, EMITNONE = 0, 1, 2, 3
LANES = np.arange(64)


def profile(cm, qdbidx, semi, L):
    lib = ia.lib()
    f = lib.infx_debug_std_profile
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_void_p, C.c_int,
                  C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p]
    cap = cm.M + 8
    buf = np.zeros((cap, 16), dtype=np.uint32)
    R, RB, nesc = C.c_int(), C.c_int(), C.c_int()
    root = (C.c_int * 9)()
    eb = C.create_string_buffer(128)
    n = f(C.cast(cm.ptr, C.c_void_p), cm.c.smx, qdbidx, semi, L, buf.ctypes.data, cap,
          C.byref(R), C.byref(RB), C.byref(nesc), root, eb)
    assert n >= 0, eb.value
    return buf[:n], R.value, RB.value, nesc.value, list(root)


def s16(x):
    x = int(x) & 0xffff
    return x - 0x10000 if x >= 0x8000 else x


class Checker:
    def __init__(self, R, RB, nesc, W, L, acc_n):
        self.R, self.RB, self.nesc, self.L, self.acc_n = R, RB, nesc, L, acc_n
        self.G = ((W + 128 + 63) // 64) * 64
        self.errors = []

    def rng(self, what, idx, lo, hi, ctx):
        idx = np.asarray(idx)
        if idx.size == 0:
            return
        mn, mx = int(idx.min()), int(idx.max())
        if mn < lo or mx >= hi:
            self.errors.append("%s index [%d, %d] outside [%d, %d) %s" % (what, mn, mx, lo, hi, ctx))

    def row(self, idx, ctx):
        self.rng("row", idx, 0, self.R, ctx)

    def rowread(self, idx, ctx):
        """A read by every lane of a block: past a row's end it lands in the
        next ring row or, after the last, in the guard of G cells."""
        self.rng("row read", idx, 0, self.R + self.G, ctx)

    def deck(self, idx, ctx):
        self.rng("deck", idx, -self.G, self.RB + self.G, ctx)

    def dsq(self, idx, ctx):
        self.rng("dsq", idx, 64 - (NR_MAX - 1), self.L + 128, ctx)

    def acc(self, idx, ctx):
        self.rng("acc", idx, 0, self.acc_n, ctx)

    def esc(self, idx, ctx):
        self.rng("esc", idx, 0, self.nesc, ctx)


def check_state(chk, w, rdn, rdx0, rdx1, W, jp0, two, j0, local):
    """Accesses of one state in the two-row sweep at chunk position jp0 (rows
    j0, j0+1 of the target, i0 = 1; residue x sits at device byte 64 + x - 1)."""
    jp1 = jp0 + 1
    dsq0 = lambda xd: 64 + j0 - xd          # left residue of cell xd, row j0
    dsq1 = lambda xd: 64 + j0 + 1 - xd      # row j0 + 1
    hdr = int(w[0])
    kind = hdr & 7
    em = (hdr >> 3) & 3
    sd = (hdr >> 5) & 3
    cnum = (hdr >> 8) & 7
    hb = local and bool((hdr >> 11) & 1)
    seq = bool((hdr >> 14) & 1)
    v = hdr >> 16
    dn = int(w[1]) & 0xffff
    sdx = int(w[1]) >> 16
    base = int(np.int32(w[2]))
    aux = int(w[3])
    dx0 = min(jp0, sdx)
    dx1 = min(jp1, sdx) if two else dn - 1
    if dn > max(dx0, dx1):
        return
    ctx = "(v %d kind %d jp0 %d two %d)" % (v, kind, jp0, two)
    bdn = max(rdn, dn)
    bdx0, bdx1 = min(rdx0, dx0), min(rdx1, dx1)
    if em != EMITNONE:
        yy = np.where(LANES < KP, LANES, 0)
        xjs = range(16) if em == EMITPAIR else [0]
        for xj in xjs:
            chk.esc(aux + (yy * KP + xj if em == EMITPAIR else yy), ctx + " ev")
    cbase = int(np.int32(w[12]))
    cb = [cbase + s16(int(w[13 + (k >> 1)]) >> (16 * (k & 1))) for k in range(6)]
    # the widest row of a visit of NR_MAX rows at this position (residue reads
    # of lanes past it are clamped to it)
    dxv = min(jp0 + NR_MAX - 1, sdx)
    if kind in (KIND_GENERIC, KIND_BEGL):
        rmask, rstride = aux & 0xffff, aux >> 16
        ir = kind == KIND_GENERIC and seq
        if ir:
            # row j+1's self term comes from row j in registers; lane 0 of the
            # first block from the stored cell dn-1 of row j
            chk.row(np.array([cb[0] + dn - 1]), ctx + " ir carry")
        for (x0, x1) in [(dx0, dx1)]:
            for d0 in range(dn, max(x0, x1) + 1, 64):
                d = d0 + LANES
                a0, a1 = d <= x0, d <= x1
                # Inside: lanes past the visit's widest band read at its edge
                # (offsets shared by the rows); CYK: idle lanes read the
                # stand-in cell dn and residue j+t
                dcl = np.minimum(d, dxv)
                chk.dsq(dsq0(dcl), ctx + " xl0")
                chk.dsq(dsq1(dcl), ctx + " xl1")
                chk.dsq(dsq1(np.where(a1, d, 1)), ctx + " xl1 cyk")
                for k in range(cnum):
                    chk.row(cb[k] + dcl - sd, ctx + " child")
                    chk.row(cb[k] + np.where(a0, d, dn) - sd, ctx + " child0 cyk")
                    chk.row(cb[k] + np.where(a1, d, dn) - sd, ctx + " child1 cyk")
                if kind == KIND_BEGL:
                    chk.deck(((jp0 & rmask) * rstride + base + d)[a0], ctx + " begl out0")
                    chk.deck(((jp1 & rmask) * rstride + base + d)[a1], ctx + " begl out1")
                else:
                    chk.row((base + d)[a0], ctx + " out0")
                    chk.row((base + d)[a1], ctx + " out1")
                    if hb:
                        chk.acc((d - rdn)[a0 & (d >= bdn) & (d <= bdx0)], ctx + " acc0")
                        chk.acc((d - rdn)[a1 & (d >= bdn) & (d <= bdx1)], ctx + " acc1")
    elif kind == KIND_IL:
        chk.row(np.array([base + dn - 1]), ctx + " carry")
        # FInside bands wider than one block take 128-cell blocks, two cells
        # per lane: every lane reads up to d0 + 127
        for d0 in range(dn, max(dx0, dx1) + 1, 64):
            for dw in ((d0 + LANES, d0 + 64 + LANES) if max(dx0, dx1) - dn >= 64 else (d0 + LANES,)):
                dcl = np.minimum(dw, dxv)
                chk.dsq(dsq0(dcl), ctx + " xl0")
                chk.dsq(dsq1(dcl), ctx + " xl1")
                for k in range(1, cnum):
                    chk.row(cb[k] + dcl - 1, ctx + " child")
            for dd in (np.where(d0 + LANES <= dx0, d0 + LANES, max(dn, dx0)), np.where(d0 + LANES <= dx1, d0 + LANES, max(dn, dx1))):
                for k in range(1, cnum):
                    chk.row(cb[k] + dd - 1, ctx + " child cyk")
            d = d0 + LANES
            a0, a1 = d <= dx0, d <= dx1
            chk.row((base + d)[a0], ctx + " out0")
            chk.row((base + d)[a1], ctx + " out1")
            if hb:
                chk.acc((d - rdn)[a0 & (d >= bdn) & (d <= bdx0)], ctx + " acc0")
                chk.acc((d - rdn)[a1 & (d >= bdn) & (d <= bdx1)], ctx + " acc1")
    else:  # B
        banded = (hdr >> 13) & 1
        kdn_y, kdx_y = int(w[6]) & 0xffff, int(w[6]) >> 16
        kdn_w, kdx_w = int(w[7]) & 0xffff, int(w[7]) >> 16
        wb = int(np.int32(w[8]))
        by = int(np.int32(w[9]))
        rmask, rstride = aux & 0xffff, aux >> 16
        for (dx, bdx) in ((dx0, bdx0), (dx1, bdx1)):
            for d0 in range(dn, dx + 1, 64):
                d = d0 + LANES
                act = d <= dx
                dd = np.where(act, d, dx)
                dl = min(d0 + 63, dx)
                if banded:
                    K0 = max(0, max(kdn_y, d0 - kdx_w))
                    K1 = min(kdx_y, dl - kdn_w)
                else:
                    K0, K1 = 0, dl
                kb = K0
                while kb <= K1:
                    ke = min(K1, kb + 63)
                    chk.row(by + np.minimum(kb + LANES, ke), ctx + " begr")
                    kk_lo, kk_hi = kb, kb + 4 * ((ke - kb) // 4) + 3    # the 4-wide steps overshoot ke
                    # any ring row may be read: bound the row part by [0, rmask * rstride]
                    chk.deck(np.array([wb - kk_hi + int(dd.min()), wb + rmask * rstride - kk_lo + int(dd.max())]),
                             ctx + " begl read")
                    kb += 64
                chk.row((base + d)[act], ctx + " out")
                if hb:
                    chk.acc((d - rdn)[act & (d >= bdn) & (d <= bdx)], ctx + " acc")


def check_root(chk, root, jp0, two, local):
    rdn, root_dx, rcnum = root[0], root[1], root[2]
    rdx0 = min(jp0, root_dx)
    rdx1 = min(jp0 + 1, root_dx) if two else rdn - 1
    for rdx in (rdx0, rdx1):
        if local:
            chk.acc(np.arange(rdn, rdx + 1) - rdn, "acc reset jp0 %d" % jp0)
        for d0 in range(rdn, rdx + 1, 64):
            d = d0 + LANES
            act = d <= rdx
            for k in range(rcnum):
                chk.row((root[3 + k] + d)[act], "root child %d jp0 %d" % (k, jp0))
            if local:
                chk.acc((d - rdn)[act], "root acc jp0 %d" % jp0)
    return rdx0, rdx1


def run_checks(name, local, qdbidx, L, semi=0):
    cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=local)
    st, R, RB, nesc, root = profile(cm, qdbidx, semi, L)
    W = min(cm.W, L)
    acc_n = root[1] - root[0] + 1 if root[0] <= root[1] else 0
    chk = Checker(R, RB, nesc, W, L, acc_n)
    rdn, root_dx = root[0], root[1]
    top = min(L, W + 2)

    def sweeps(points):
        """(jp0, two) pairs of the first chunk (i0 = 1, so j0 = jp0) at the given positions"""
        out = set()
        for p in points:
            if 1 <= p <= top:
                out.add((p, False))
                if p + 1 <= L:
                    out.add((p, True))
        return sorted(out)

    # the root band, the accumulators and every state at the positions where
    # their clips change (indices depend on jp only through min(jp, dx))
    for jp0, two in sweeps([1, 2, 3, rdn - 1, rdn, root_dx - 1, root_dx, W - 1, W, W + 1]):
        check_root(chk, root, jp0, two, local)
    for w in st:
        dn, sdx = int(w[1]) & 0xffff, int(w[1]) >> 16
        for jp0, two in sweeps([1, 2, dn - 2, dn - 1, dn, dn + 1, sdx - 2, sdx - 1, sdx, sdx + 1, W - 1, W, W + 1]):
            rdx0 = min(jp0, root_dx)
            rdx1 = min(jp0 + 1, root_dx) if two else rdn - 1
            check_state(chk, w, rdn, rdx0, rdx1, W, jp0, two, jp0, local)
    # a visit of NR rows reads the residue j0+t of each idle row t (past the
    # chunk's end: up to j0+NR-1 with j0 <= L), on every lane
    chk.dsq(np.array([64 + L + NR_MAX - 2]), "idle-row residue at the target's end")
    # the target's end (last chunk, steady-state bands): a row pair ending at L and a lone row at L
    for j0, two in ((L - 1, True), (L, False)):
        if j0 < 1:
            continue
        jp_end = min(j0, W + 2)          # a chunk starts at or after residue 1
        rdx0 = min(jp_end, root_dx)
        rdx1 = min(jp_end + 1, root_dx) if two else rdn - 1
        for w in st:
            check_state(chk, w, rdn, rdx0, rdx1, W, jp_end, two, j0, local)
    assert not chk.errors, "%s local=%s qdb=%d L=%d: %d bad accesses, first: %s" % (
        name, local, qdbidx, L, len(chk.errors), chk.errors[:5])


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Vault.c.cm", "5S_rRNA.c.cm", "Cobalamin.c.cm", "Plant_SRP.c.cm",
                                  "rnaseP-eubact.c.cm"])
def test_std_kernel_addresses_in_bounds(name):
    cm = ia.CM.read(os.path.join(CMDIR, name))
    W = cm.W
    small = name in ("tRNA.c.cm", "5S_rRNA.c.cm")
    for local in (True, False):
        for qdbidx in ((0, 1, 2) if small else (1, 2)):
            lengths = [1, 2, 7, 64, W - 1, W, W + 1, 3 * W] if small else [W - 1, 3 * W]
            for L in lengths:
                run_checks(name, local, qdbidx, L)


@pytest.mark.parametrize("name", ["tRNA.c.cm", "Cobalamin.c.cm", "rnaseP-eubact.c.cm", "synth_multiloop350.cm",
                                  "synth_nested1500.cm"])
def test_begl_rings_hold_one_visit(name):
    """A state visit writes rows jp .. jp+NR-1 of a BEGL deck before its B
    parent reads deck rows jp+t-k, k up to the largest split maxk: the ring
    (rmask + 1 rows) must keep rows jp-maxk .. jp+NR-1 apart, i.e. hold
    maxk + NR rows, for every bifurcation, band set and mode."""
    for local in (True, False):
        cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=local)
        for qdbidx in (0, 1, 2):
            st = profile(cm, qdbidx, 0, 100000)[0]
            nb = 0
            for w in st:
                hdr = int(w[0])
                if hdr & 7 != KIND_B:
                    continue
                nb += 1
                dn, sdx = int(w[1]) & 0xffff, int(w[1]) >> 16
                kdn_y, kdx_y = int(w[6]) & 0xffff, int(w[6]) >> 16
                kdn_w, kdx_w = int(w[7]) & 0xffff, int(w[7]) >> 16
                if (hdr >> 13) & 1:
                    maxk = max((min(kdx_y, d - kdn_w) for d in range(dn, sdx + 1)
                                if max(0, kdn_y, d - kdx_w) <= min(kdx_y, d - kdn_w)), default=0)
                else:
                    maxk = sdx
                ring = (int(w[3]) & 0xffff) + 1
                assert ring & (ring - 1) == 0
                assert ring >= maxk + NR_MAX, "%s qdb %d state %d: ring %d, maxk %d" % (name, qdbidx, hdr >> 16, ring, maxk)
            assert nb > 0


def test_checker_catches_planted_overreach():
    """The checker itself: on the first row of a chunk at the target's start
    (j0 = 1), a lane reading the left residue of d = NR_MAX + 1 (target
    position 1 - NR_MAX, NR_MAX bytes before the target) is reported, while
    the NR_MAX - 1 bytes the clamped reads may touch are not; a row read one
    cell past the guard is reported."""
    chk = Checker(R=10, RB=10, nesc=10, W=5, L=5, acc_n=3)
    j0 = 1
    chk.dsq(np.array([64 + j0 - 1, 64 + j0 - NR_MAX]), "residue j0, and NR_MAX - 1 before it")
    chk.rowread(np.array([0, 10 + chk.G - 1]), "row reads up to the guard's end")
    assert not chk.errors
    chk.dsq(np.array([64 + j0 - 1 - NR_MAX]), "planted")
    assert len(chk.errors) == 1
    chk.rowread(np.array([10 + chk.G]), "planted")
    assert len(chk.errors) == 2


def _blocks(dn, dxm, nbw):
    """Block starts gen_state2 (scan_std.hip) visits with up to nbw blocks per step."""
    out, d0 = [], dn
    if nbw >= 4:
        while d0 + 192 <= dxm:
            out += [d0, d0 + 64, d0 + 128, d0 + 192]
            d0 += 256
    if nbw >= 2:
        while d0 + 64 <= dxm:
            out += [d0, d0 + 64]
            d0 += 128
    while d0 <= dxm:
        out.append(d0)
        d0 += 64
    return out


def test_multi_block_steps_visit_the_same_blocks():
    for dn in (0, 1, 2, 5, 63, 64):
        for dxm in range(dn - 1, dn + 700):
            one = _blocks(dn, dxm, 1)
            assert one == list(range(dn, dxm + 1, 64))
            assert _blocks(dn, dxm, 2) == one and _blocks(dn, dxm, 4) == one, (dn, dxm)
