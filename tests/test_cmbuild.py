"""Model construction (infernal_amd.cmbuild).

Topology is pinned against the models the reference's cmbuild wrote: the
consensus structure of each shipped model is read back from its node lines
(MATP pairs, MATL/MATR singles, positions ranked into consensus order) and
the guide tree, state layout and CMRebalance numbering built from it must
reproduce the file's node types and every state's type, cfirst, cnum, plast
and pnum.

Parameters are pinned against the three tutorial models that the current
cmbuild wrote with default options from alignments the reference ships
(tutorial/tRNA5.sto, 5S_rRNA.sto, Cobalamin.sto -> the committed
tRNA5.c.cm, 5S_rRNA.c.cm, Cobalamin.c.cm): the rebuilt model must have the
same CLEN, STATES, W and EFFN, every transition and emission score within
the files' print rounding (3 decimals), and the same bands.  The alignments
and the default prior file are read from the reference tree, so these tests
skip where it is absent.  (The testsuite models were written by a 2012
development build with other weighting defaults and are not used here.)
"""
import os

import numpy as np
import pytest

from infernal_amd import cmbuild as cb

CMDIR = os.path.join(os.path.dirname(__file__), "golden", "cm")
MODELS = ["tRNA.c.cm", "Vault.c.cm", "5S_rRNA.c.cm", "Cobalamin.c.cm", "Plant_SRP.c.cm", "rnaseP-eubact.c.cm", "tRNA5.c.cm"]
NODE = {"BIF": cb.BIF_nd, "MATP": cb.MATP_nd, "MATL": cb.MATL_nd, "MATR": cb.MATR_nd, "BEGL": cb.BEGL_nd,
        "BEGR": cb.BEGR_nd, "ROOT": cb.ROOT_nd, "END": cb.END_nd}
STATE = {v: k for k, v in cb.STATE_NAMES.items()}


def read_layout(path):
    nodes, pos, pairs, stl = [], [], [], []
    in_cm = False
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "CM":
            in_cm = True
            continue
        if not in_cm or t[0] == "//":
            continue
        if t[0] == "[":
            nt = NODE[t[1]]
            nodes.append(nt)
            l, r = t[4], t[5]
            if nt == cb.MATP_nd:
                pos += [int(l), int(r)]; pairs.append((int(l), int(r)))
            elif nt == cb.MATL_nd:
                pos.append(int(l))
            elif nt == cb.MATR_nd:
                pos.append(int(r))
            continue
        stl.append((STATE[t[0]], int(t[4]), int(t[5]), int(t[2]), int(t[3])))
    rank = {p: k + 1 for k, p in enumerate(sorted(pos))}
    ct = [0] * (len(pos) + 1)
    for l, r in pairs:
        ct[rank[l]], ct[rank[r]] = rank[r], rank[l]
    return nodes, stl, ct


@pytest.mark.parametrize("name", MODELS)
def test_topology_matches_shipped_models(name):
    nodes, stl, ct = read_layout(os.path.join(CMDIR, name))
    nd_types, nodemap, st, ndidx, cfirst, cnum, plast, pnum = cb.layout_for_structure(ct)
    in_state_order = [nd_types[n] for n in sorted(range(len(nd_types)), key=lambda n: nodemap[n])]
    assert in_state_order == nodes
    got = list(zip(st, cfirst, cnum, plast, pnum))
    assert len(got) == len(stl)
    bad = [(v, got[v], stl[v]) for v in range(len(stl)) if got[v] != stl[v]]
    assert not bad, bad[:5]


def test_ss_to_ct_wuss():
    ct = cb.ss_to_ct("<<..[[..]]..>>")
    assert ct[1] == 14 and ct[2] == 13 and ct[5] == 10 and ct[6] == 9 and ct[3] == 0


REF = os.environ.get("INFERNAL_REFERENCE", "/root/reference")
PRIOR = os.path.join(REF, "src", "default.prior")
need_ref = pytest.mark.skipif(not os.path.exists(PRIOR), reason="reference tree (alignments, default prior) not present")


def _parse_cm(text):
    hdr, states = {}, []
    in_cm = False
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        if not in_cm:
            if f[0] == "CM":
                in_cm = True
            else:
                hdr.setdefault(f[0], f[1:])
            continue
        if f[0] == "//":
            break
        if f[0] != "[":
            states.append(f)
    return hdr, states


@need_ref
@pytest.mark.parametrize("name", ["tRNA5", "5S_rRNA", "Cobalamin"])
def test_build_reproduces_tutorial_models(name):
    aseqs, gc = cb.read_stockholm(os.path.join(REF, "tutorial", name + ".sto"))
    ax = cb.digitize_alignment(aseqs)
    text, info = cb.build_cm_from_alignment(ax, gc["SS_cons"], name, PRIOR)
    h1, s1 = _parse_cm(text)
    h2, s2 = _parse_cm(open(os.path.join(CMDIR, name + ".c.cm")).read())
    for key in ("CLEN", "STATES", "NODES", "W", "NSEQ"):
        assert h1[key] == h2[key], key
    assert "%.6f" % float(h1["EFFN"][0]) == h2["EFFN"][0]
    assert len(s1) == len(s2)
    band_rows = 0
    for a, b in zip(s1, s2):
        assert a[:6] == b[:6]
        ba, bb = [int(x) for x in a[6:10]], [int(x) for x in b[6:10]]
        band_rows += ba != bb
        assert max(abs(x - y) for x, y in zip(ba, bb)) <= 1
        va = [-1e30 if x == "*" else float(x) for x in a[10:]]
        vb = [-1e30 if x == "*" else float(x) for x in b[10:]]
        assert len(va) == len(vb)
        assert all(abs(x - y) <= 0.0006 for x, y in zip(va, vb)), (a[:2], a[10:], b[10:])
    assert band_rows <= 1


@need_ref
@pytest.mark.parametrize("name", ["tRNA5", "5S_rRNA", "Cobalamin"])
def test_bands_of_rebuilt_models_exact_but_for_ties(name):
    """The band engine on the models rebuilt from the tutorial alignments,
    whose transition scores are written to six decimals (cmbuild's own
    bands came from its unrounded probabilities, the shipped files print
    three): W and every band column equal the shipped file's, except where
    the tail mass at the contested length is within 1e-3 of beta (one column
    of the three models: Cobalamin state 40's dmax2, 7.8e-5 from beta2)."""
    import infernal_amd as ia
    from test_qdband import contested, engine_bands, qdb_bands
    aseqs, gc = cb.read_stockholm(os.path.join(REF, "tutorial", name + ".sto"))
    text, info = cb.build_cm_from_alignment(cb.digitize_alignment(aseqs), gc["SS_cons"], name, PRIOR)
    shipped = ia.CM.read(os.path.join(CMDIR, name + ".c.cm"))
    cm = ia.CM.from_text(text)
    got, W, Z, gamma = engine_bands(cm, cm.c.beta_W)
    assert W == shipped.W
    q = cm.c.qdbinfo.contents
    ties = contested(got, qdb_bands(shipped), gamma, q.beta1, q.beta2)
    assert len(ties) <= 1
    for nm, v, f, o, n, ratio in ties:
        assert abs(f - o) == 1 and abs(ratio - 1.0) <= 1e-3, (nm, v, f, o, ratio)


@need_ref
def test_pb_weights_and_match_columns():
    """Weights sum to nseq; an all-identical alignment weighs sequences
    equally; a gap-majority column is an insert column."""
    ax = cb.digitize_alignment(["ACGU-A", "ACGU-A", "ACGUCA"])
    w = cb.pb_weights(ax)
    assert abs(w.sum() - 3) < 1e-12 and np.allclose(w, 1.0)
    mat = cb.match_columns(ax, w)
    assert mat[1:7] == [True, True, True, True, False, True]
    ax = cb.digitize_alignment(["AAAA", "AAAC", "CCCC"])
    w = cb.pb_weights(ax)
    # column 4: A once, C twice; columns 1-3: A twice, C once
    raw = np.array([3 * 1 / 4 + 1 / 2, 3 * 1 / 4 + 1 / 4, 3 * 1 / 2 + 1 / 4]) / 4
    assert np.allclose(w, raw * 3 / raw.sum())
