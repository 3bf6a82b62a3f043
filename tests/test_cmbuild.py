"""Model construction (infernal_amd.cmbuild).

Topology is pinned against the models the reference's cmbuild wrote: the
consensus structure of each shipped model is read back from its node lines
(MATP pairs, MATL/MATR singles, positions ranked into consensus order) and
the guide tree, state layout and CMRebalance numbering built from it must
reproduce the file's node types and every state's type, cfirst, cnum, plast
and pnum.
"""
import os

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
