"""Host-side model input and configuration, pinned against the .cm files'
own printed scores and against closed forms (no GPU needed)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import infernal_amd as ia
from helpers import CMDIR, ROOT

ALL_CMS = sorted(f for f in os.listdir(CMDIR) if f.endswith(".cm"))


def _file_state_lines(path):
    """(type, fields) of every state line, in order."""
    out = []
    in_cm = False
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if t[0] == "CM":
                in_cm = True
                continue
            if not in_cm or t[0] == "[":
                continue
            if t[0] == "//":
                break
            out.append(t)
    return out


@pytest.mark.parametrize("name", ALL_CMS)
def test_read_matches_file(name):
    """Global-mode log-odds scores reproduce the file's %.3f values
    (src/cm_file.c:ascii2prob, src/cm.c:CMRenormalize and CMLogoddsify)."""
    path = os.path.join(CMDIR, name)
    cm = ia.CM.read(path)
    cm.configure(local=False)
    c = cm.c
    lines = _file_state_lines(path)
    assert len(lines) == c.M
    nullv = [c.null[x] for x in range(4)]
    for v, t in enumerate(lines):
        st = ord(c.sttype[v])
        assert int(t[1]) == v
        assert c.cfirst[v] == int(t[4]) and c.cnum[v] == int(t[5])
        q = c.qdbinfo.contents
        assert (q.dmin2[v], q.dmin1[v], q.dmax1[v], q.dmax2[v]) == tuple(int(x) for x in t[6:10])
        k = 10
        if st != ia.B_st:
            for x in range(c.cnum[v]):
                if t[k] != "*":
                    assert abs(c.tsc[v][x] - float(t[k])) < 6e-4, (v, x, c.tsc[v][x], t[k])
                k += 1
        if st in (ia.ML_st, ia.MR_st, ia.IL_st, ia.IR_st):
            for x in range(4):
                assert abs(c.esc[v][x] - float(t[k + x])) < 6e-4
                assert abs(c.oesc[v][x] - float(t[k + x])) < 6e-4
            # degenerate N = average of the four scores (esl_abc_FAvgScore)
            assert abs(c.oesc[v][15] - sum(c.esc[v][x] for x in range(4)) / 4) < 1e-6
            assert c.oesc[v][4] < -9e35 and c.oesc[v][17] < -9e35
        if st == ia.MP_st:
            for x in range(16):
                assert abs(c.esc[v][x] - float(t[k + x])) < 6e-4
    # background renormalized
    assert abs(sum(nullv) - 1.0) < 1e-6


def test_integer_scores_round_float_scores():
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    c = cm.c
    for v in range(c.M):
        st = ord(c.sttype[v])
        if st in (ia.B_st, ia.E_st):
            continue
        for x in range(c.cnum[v]):
            f = c.tsc[v][x]
            if f > -1e30:
                assert abs(c.itsc[v][x] - f * 1000) <= 1.0
        if c.beginsc[v] > -1e30:
            assert abs(c.ibeginsc[v] - 1000 * c.beginsc[v]) <= 1.0
        else:
            assert c.ibeginsc[v] == -987654321


def test_localize_closed_forms():
    """cm_localize (src/cm_modelconfig.c:467-548): begin mass 1-pbegin on
    node 1, pbegin spread evenly over internal MATP/MATL/MATR/BIF nodes; end
    mass pend spread over exit nodes; root transitions zeroed."""
    path = os.path.join(CMDIR, "Vault.c.cm")
    g = ia.CM.read(path).configure(local=False)
    l = ia.CM.read(path).configure(local=True)
    cg, cl = g.c, l.c
    ndtype = [ord(cl.ndtype[n]) for n in range(cl.nodes)]
    nodemap = [cl.nodemap[n] for n in range(cl.nodes)]
    starts = [n for n in range(2, cl.nodes) if ndtype[n] in (0, 1, 2, 3)]
    exits = [n for n in range(1, cl.nodes) if ndtype[n] in (1, 2, 3, 4, 5) and n + 1 < cl.nodes and ndtype[n + 1] != 7]
    assert abs(cl.begin[nodemap[1]] - 0.95) < 1e-7
    for n in starts:
        assert abs(cl.begin[nodemap[n]] - 0.05 / len(starts)) < 1e-9
    assert abs(sum(cl.begin[v] for v in range(cl.M)) - 1.0) < 1e-5
    exit_states = {nodemap[n] for n in exits}
    for v in range(cl.M):
        if v in exit_states:
            assert abs(cl.end[v] - 0.05 / len(exits)) < 1e-9
            tot = sum(cl.t[v][k] for k in range(cl.cnum[v])) + cl.end[v]
            assert abs(tot - 1.0) < 1e-5
            for k in range(cl.cnum[v]):
                assert abs(cl.t[v][k] - cg.t[v][k] / (1 + cl.end[v])) < 1e-6
        else:
            assert cl.end[v] == 0.0
    for k in range(cl.cnum[0]):
        assert cl.tsc[0][k] < -9e35
    assert cl.flags & ia.CMH_LOCAL_BEGIN and cl.flags & ia.CMH_LOCAL_END


def test_begin_end_scores_closed_forms():
    """CMLogoddsify's begin/end scores (src/cm.c:868-880): beginsc = log2
    begin, endsc = log2 end (IMPOSSIBLE where the probability is 0), their
    integer twins Prob2Score(p, 1) = floor(0.5 + 1000 log2 p) (-INFTY at 0),
    and iel_selfsc = Prob2Score(2^el_selfsc, 1), on the localised model whose
    probabilities test_localize_closed_forms pins."""
    import math
    for name in ("Vault.c.cm", "Cobalamin.c.cm"):
        cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=True)
        c = cm.c
        n_on = 0
        for v in range(c.M):
            for prob, sc, isc in ((c.begin[v], c.beginsc[v], c.ibeginsc[v]), (c.end[v], c.endsc[v], c.iendsc[v])):
                if prob == 0.0:
                    assert sc < -9e35 and isc <= -987654321, (name, v)
                    continue
                lg = math.log(prob) * 1.44269504
                assert abs(sc - lg) <= 1e-6 * max(1.0, abs(lg)), (name, v, sc, lg)
                assert isc == math.floor(0.5 + 1000.0 * lg), (name, v, isc, lg)
                n_on += 1
        assert n_on > 50
        assert c.iel_selfsc == math.floor(0.5 + 1000.0 * math.log(2.0 ** c.el_selfsc) * 1.44269504)


def test_logsum_tables_closed_form():
    """FLogsum table entry i = log2(1 + 2^(-i/1000)); ILogsum table its x1000
    rounding (src/logsum.c:74-148).  The reference evaluates these with the
    truncated constants 1.44269504 and 0.69314718, so allow 1e-6 bits."""
    lib = ia.lib()
    for i in (0, 1, 2, 17, 999, 1000, 4321, 12345, 22999):
        exact = math.log2(1 + 2 ** (-i / 1000))
        a, b = 5.0, 5.0 - i / 1000.0
        got = lib.FLogsum(a, b) - a
        # index = (int)((max-min)*1000) may land one entry lower when the float
        # difference rounds below i/1000
        assert abs(got - exact) < 2e-3
        assert abs(lib.ILogsum(5000, 5000 - i) - 5000 - round(1000 * exact)) <= 1
    assert lib.FLogsum(3.0, 30.0) == 30.0          # difference >= 23: max
    assert lib.FLogsum(-1e36, 2.0) == 2.0
    assert abs(lib.FLogsum(1.5, 1.5) - 2.5) < 1e-6  # table[0] = 1
    assert lib.ILogsum(-987654321, 100) == 100
    assert lib.ILogsum(100, 100) == 1100


def test_abi_exports_every_declared_function():
    hdr = open(os.path.join(ROOT, "include", "infernal_gpu.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"^\s*(?:int|void|float|CM_TOPHITS \*)\s+\**([A-Za-z_]\w*)\s*\(", hdr, flags=re.M))
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", ia.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing


def test_contract_errors():
    """Contract violations return eslEINCOMPAT before any device work
    (src/cm_dpsearch.c:129-135, :1771-1776)."""
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    dsq = ia.dsq_from_codes(np.zeros(50, dtype=np.uint8))
    with pytest.raises(ia.ScanError) as e:
        ia.FastCYKScan(cm, 1, dsq, 10, 5, 0.0)
    assert e.value.status == ia.eslEINCOMPAT
    with pytest.raises(ia.ScanError) as e:
        ia.FastCYKScan(cm, 3, dsq, 1, 50, 0.0)
    assert e.value.status == ia.eslEINCOMPAT
    # Inside scanners require CM_SEARCH_INSIDE; the wrapper sets it, so call the C ABI directly
    import ctypes as C
    lib = ia.lib()
    cm.c.search_opts = 0
    eb = C.create_string_buffer(128)
    st = lib.FastFInsideScan(cm.ptr, eb, cm.c.smx, 1, dsq.ctypes.data, 1, 50, 0.0, None, 0, 0.0, None, None, None, None)
    assert st == ia.eslEINCOMPAT and b"CM_SEARCH_INSIDE" in eb.value
    bad = dsq.copy()
    bad[5] = 17
    with pytest.raises(ia.ScanError) as e:
        ia.FastCYKScan(cm, 1, bad, 1, 50, 0.0)
    assert e.value.status == ia.eslEINVAL


def test_positions_past_32_bits_are_refused():
    """Hit candidates carry 32-bit positions: a call whose j0 passes
    2^31 - 1 is refused with eslERANGE before the residues are read (the
    residue pointer given here is never dereferenced), by the standard and
    the truncated scanners alike.  cmsearch scans buffers of one sequence
    window, so its positions stay far below that."""
    import ctypes as C
    from helpers import load_cm
    lib = ia.lib()
    eb = C.create_string_buffer(128)
    i0, j0 = 2 ** 31 - 100, 2 ** 31 + 100
    fake = C.c_void_p(4096)                     # never read: the range check comes first
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    st = lib.FastCYKScan(cm.ptr, eb, cm.c.smx, 1, fake, i0, j0, 0.0, None, 0, 0.0, None, None, None, None)
    assert st == 16 and b"do not fit the hit candidates" in eb.value, eb.value   # eslERANGE
    cmt = load_cm("tRNA.c.cm", local=True, trunc=True)
    st = lib.RefTrCYKScan(cmt.ptr, eb, cmt.c.trsmx, 1, 2, fake, i0, j0, 0.0, None, 0, 0.0, None, None, None, None, None)
    assert st == 16 and b"do not fit the hit candidates" in eb.value, eb.value


def test_windowed_entry_points_contract_errors():
    """infx_FastScanWindows and infx_TrScanWindows check their windows
    before any device work: sorted and disjoint (standard), j0 >= i0 and
    i0 >= 1 (truncated), residue codes, the pass index; without a GPU a
    valid call reports eslENODEVICE."""
    from helpers import load_cm
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    dsq = ia.dsq_from_codes(np.zeros(600, dtype=np.uint8))
    with pytest.raises(ia.ScanError) as e:
        ia.ScanWindows("cyk", cm, 1, dsq, [(1, 100), (50, 150)], 0.0)
    assert e.value.status == ia.eslEINCOMPAT and "sorted and disjoint" in str(e.value)
    with pytest.raises(ia.ScanError) as e:
        ia.ScanWindows("cyk", cm, 1, dsq, [(200, 300), (1, 100)], 0.0)
    assert e.value.status == ia.eslEINCOMPAT
    bad = dsq.copy()
    bad[250] = 17
    with pytest.raises(ia.ScanError) as e:
        ia.ScanWindows("finside", cm, 2, bad, [(1, 100), (200, 300)], 0.0)
    assert e.value.status == ia.eslEINVAL
    cmt = load_cm("tRNA.c.cm", local=True, trunc=True)
    with pytest.raises(ia.ScanError) as e:
        ia.TrScanWindows("cyk", cmt, 1, 2, dsq, [(100, 50)], 0.0)
    assert e.value.status == ia.eslEINCOMPAT
    with pytest.raises(ia.ScanError) as e:
        ia.TrScanWindows("cyk", cmt, 1, 99, dsq, [(1, 100)], 0.0)
    assert e.value.status in (ia.eslEINVAL, ia.eslEINCOMPAT)
    with pytest.raises(ia.ScanError) as e:
        ia.TrScanWindows("iinside", cmt, 2, 2, bad, [(200, 300)], 0.0)
    assert e.value.status == ia.eslEINVAL
    if not ia.gpu_available():
        with pytest.raises(ia.ScanError) as e:
            ia.ScanWindows("cyk", cm, 1, dsq, [(1, 100), (200, 300)], 0.0)
        assert e.value.status == ia.eslENODEVICE
        with pytest.raises(ia.ScanError) as e:
            ia.TrScanWindows("cyk", cmt, 1, 2, dsq, [(1, 100), (200, 300)], 0.0)
        assert e.value.status == ia.eslENODEVICE


def test_batch_contract_errors():
    """infx_FastScanBatch checks its windows and every model before any
    device work: model indices in range, each model's windows sorted and
    disjoint, the Inside flag, residue codes."""
    cms = [ia.CM.read(os.path.join(CMDIR, n)).configure(local=True) for n in ("tRNA.c.cm", "Vault.c.cm")]
    dsq = ia.dsq_from_codes(np.zeros(400, dtype=np.uint8))
    cut = [0.0, 0.0]
    with pytest.raises(ia.ScanError) as e:
        ia.ScanBatch("cyk", cms, 1, dsq, [(1, 100, 0), (50, 150, 2)], cut)
    assert e.value.status == ia.eslEINVAL
    with pytest.raises(ia.ScanError) as e:
        ia.ScanBatch("cyk", cms, 1, dsq, [(1, 100, 0), (50, 150, 0)], cut)
    assert e.value.status == ia.eslEINCOMPAT and "sorted and disjoint" in str(e.value)
    bad = dsq.copy()
    bad[120] = 17
    with pytest.raises(ia.ScanError) as e:
        ia.ScanBatch("cyk", cms, 1, bad, [(1, 100, 0), (50, 150, 1)], cut)
    assert e.value.status == ia.eslEINVAL
    if not ia.gpu_available():
        with pytest.raises(ia.ScanError) as e:      # windows of different models may overlap
            ia.ScanBatch("finside", cms, 2, dsq, [(1, 100, 0), (50, 150, 1), (200, 390, 0)], cut)
        assert e.value.status == ia.eslENODEVICE


@pytest.mark.skipif(ia.gpu_available(), reason="only meaningful without a GPU")
def test_no_gpu_fails_loudly():
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    dsq = ia.dsq_from_codes(np.zeros(50, dtype=np.uint8))
    with pytest.raises(ia.ScanError) as e:
        ia.FastCYKScan(cm, 1, dsq, 1, 50, 0.0)
    assert e.value.status == ia.eslENODEVICE


def test_cells_per_residue_match_shipped_model_table():
    """Cells and ops per residue at steady state for the shipped models
    (QDB1, beta=1e-7): the figures measured independently from the .cm band
    columns for the benchmark write-up."""
    expect = {"tRNA.c.cm": (14629, 162506), "Vault.c.cm": (34899, 315066), "5S_rRNA.c.cm": (16134, 163376),
              "Cobalamin.c.cm": (86217, 863089), "Plant_SRP.c.cm": (70370, 810985),
              "rnaseP-eubact.c.cm": (83094, 841718)}
    for name, (cells, ops) in expect.items():
        cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=True)
        c1, o1 = cm.cells_per_res(1)
        assert (c1, o1) == (cells, ops), (name, c1, o1)


@pytest.mark.skipif(ia.gpu_available(), reason="without a GPU the scanners stop right after building their device profile")
@pytest.mark.parametrize("name", ["tRNA.c.cm", "Vault.c.cm", "5S_rRNA.c.cm", "Cobalamin.c.cm", "Plant_SRP.c.cm",
                                  "rnaseP-eubact.c.cm", "tRNA5.c.cm"])
def test_scan_profiles_build_for_every_model_and_length(name):
    """The host-side scan profiles (packed state descriptors, deck rings,
    truncated penalty lists) build for every shipped model, both modes, all
    band sets and target lengths around W: the only error left is the
    missing device."""
    rng = np.random.default_rng(0)
    for local in (True, False):
        cm = ia.CM.read(os.path.join(CMDIR, name)).configure(local=local, trunc=True, trscanmx=True)
        W = cm.W
        for L in (1, 2, 7, W - 1, W, W + 1, 3 * W):
            if L < 1:
                continue
            dsq = ia.dsq_from_codes(rng.integers(0, 4, L).astype(np.uint8))
            for fn in (ia.FastCYKScan, ia.FastFInsideScan, ia.FastIInsideScan):
                for q in (0, 1, 2):
                    with pytest.raises(ia.ScanError) as e:
                        fn(cm, q, dsq, 1, L, 0.0)
                    assert e.value.status == ia.eslENODEVICE, (name, local, L, fn.__name__, q, str(e.value))
            for fn in (ia.RefTrCYKScan, ia.RefITrInsideScan):
                for p in (ia.PLI_PASS_5P_ONLY_FORCE, ia.PLI_PASS_5P_AND_3P_ANY):
                    with pytest.raises(ia.ScanError) as e:
                        fn(cm, 1, p, dsq, 1, L, 0.0)
                    assert e.value.status == ia.eslENODEVICE, (name, local, L, fn.__name__, p, str(e.value))


# IUPAC-IUB nucleotide codes (Cornish-Bowden 1985, Nucleic Acids Res 13:3021), U for T
IUPAC = {"A": "A", "C": "C", "G": "G", "U": "U", "T": "U", "R": "AG", "Y": "CU", "M": "AC", "K": "GU",
         "S": "CG", "W": "AU", "H": "ACU", "B": "CGU", "V": "ACG", "D": "AGU", "N": "ACGU", "X": "ACGU"}


def test_degenerate_residue_scores_follow_iupac():
    """Every IUPAC letter, upper or lower case, digitises to the code whose
    optimised emission score is the mean of the canonical scores of the
    residues the letter stands for: for singlet states, and for both sides
    of a pair (esl_abc_FAvgScore, src/cm.c CalcOptimizedEmitScores).  The
    gap code scores IMPOSSIBLE ('*' and '~', codes above 15, are refused by
    the scanners).  A swapped row of the degeneracy table or a wrong symbol
    order fails here."""
    cm = ia.CM.read(os.path.join(CMDIR, "tRNA.c.cm")).configure(local=True)
    c = cm.c
    Kp = 18
    code = {ch: int(ia.digitize(ch)[1]) for ch in IUPAC}
    for ch in IUPAC:
        assert int(ia.digitize(ch.lower())[1]) == code[ch]
    nuc = {ch: ["ACGU".index(y) for y in IUPAC[ch]] for ch in IUPAC}
    codes_seen = set(code.values())
    assert len(codes_seen) == 15          # X shares N's code, T shares U's
    nsing = npair = 0
    for v in range(c.M):
        st = ord(c.sttype[v])
        if st in (ia.ML_st, ia.MR_st, ia.IL_st, ia.IR_st):
            nsing += 1
            for ch, ys in nuc.items():
                want = sum(c.esc[v][y] for y in ys) / len(ys)
                assert abs(c.oesc[v][code[ch]] - want) < 1e-5, (v, ch)
            assert c.oesc[v][4] < -9e35
        elif st == ia.MP_st:
            npair += 1
            for a, la in nuc.items():
                for b, rb in nuc.items():
                    want = sum(c.esc[v][l * 4 + r] for l in la for r in rb) / (len(la) * len(rb))
                    assert abs(c.oesc[v][code[a] * Kp + code[b]] - want) < 1e-4, (v, a, b)
            assert all(c.oesc[v][4 * Kp + b] < -9e35 and c.oesc[v][b * Kp + 4] < -9e35 for b in range(16))
    assert nsing > 10 and npair > 10


def test_marginal_emission_scores(tmp_path):
    """The truncated scanners' marginal scores (src/cm.c CMLogoddsify): for a
    pair state the left marginal of residue x is the log-odds of the pair
    probabilities summed over the right residue (and the right marginal over
    the left one), recovered here from the pair scores; a degenerate code
    scores the background-weighted mean over the residues its IUPAC letter
    stands for (esl_abc_FExpectScVec).  Singlet states carry their own scores
    on their side and 0 on the other.  The model's background is replaced by
    a skewed one so that the weighting shows."""
    with open(os.path.join(CMDIR, "Cobalamin.c.cm")) as f:
        text = f.read()
    text, n = re.subn(r"^NULL .*$", "NULL     0.500 -0.700  0.200 -0.300 ", text, count=1, flags=re.M)
    assert n == 1
    path = tmp_path / "Cobalamin.skewed.cm"
    path.write_text(text)
    cm = ia.CM.read(str(path)).configure(local=True, trunc=True)
    c = cm.c
    null = [c.null[x] for x in range(4)]
    # the background as read: 0.25 * 2^x, renormalised (esl_vec_FNorm)
    raw = [0.25 * 2.0 ** x for x in (0.5, -0.7, 0.2, -0.3)]
    assert all(abs(null[x] - raw[x] / sum(raw)) < 1e-6 for x in range(4))
    code = {ch: int(ia.digitize(ch)[1]) for ch in IUPAC}
    nuc = {ch: ["ACGU".index(y) for y in IUPAC[ch]] for ch in IUPAC}

    def expect(sc, ys):
        return sum(null[y] * sc[y] for y in ys) / sum(null[y] for y in ys)

    seen = set()
    for v in range(c.M):
        st = ord(c.sttype[v])
        lm, rm = c.lmesc[v], c.rmesc[v]
        if st == ia.MP_st:
            for x in range(4):
                pl = sum(2.0 ** c.esc[v][x * 4 + y] * null[x] * null[y] for y in range(4))
                pr = sum(2.0 ** c.esc[v][y * 4 + x] * null[y] * null[x] for y in range(4))
                assert abs(lm[x] - math.log2(pl / null[x])) < 1e-4, (v, x)
                assert abs(rm[x] - math.log2(pr / null[x])) < 1e-4, (v, x)
            sides = (lm, rm)
        elif st in (ia.ML_st, ia.IL_st):
            assert all(abs(lm[x] - c.esc[v][x]) < 1e-6 and rm[x] == 0.0 for x in range(4))
            sides = (lm,)
        elif st in (ia.MR_st, ia.IR_st):
            assert all(abs(rm[x] - c.esc[v][x]) < 1e-6 and lm[x] == 0.0 for x in range(4))
            sides = (rm,)
        else:
            continue
        seen.add(st)
        for sc in sides:
            for ch, ys in nuc.items():
                assert abs(sc[code[ch]] - expect(sc, ys)) < 1e-5, (v, ch)
    assert {ia.MP_st, ia.ML_st, ia.MR_st, ia.IL_st, ia.IR_st} <= seen
