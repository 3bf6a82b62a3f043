"""Extract the hand-enumerated truncated alignment paths of the reference's
brute-force integration test into a JSON fixture.

src/itest_brute.c (score_brute_matl_cm) lists, for a ROOT-MATL-MATL-END
model, every Joint, Left-marginal and Right-marginal alignment of a one- and
a two-residue sequence as a product of named parameters, e.g.
    Lp[4] = msc * prm->t2t3;   (ROOT_IR(2L) MATL_ML(3L))
This script keeps those products as data: per mode and sequence length, the
list of factor names.  The length of a path is the number of residues it
emits: one per msc / isc factor and one per EL self-transition (el_self).

Usage:  python tests/golden/make_itest_brute_paths.py [/root/reference]
writes tests/golden/itest_brute_matl_paths.json
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "src", "itest_brute.c")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "itest_brute_matl_paths.json")


def factor(tok):
    tok = tok.strip()
    tok = tok.replace("prm->", "")
    m = re.fullmatch(r"(end|begin)\[(\d+)\]", tok)
    if m:
        return m.group(1) + m.group(2)
    assert re.fullmatch(r"t\d+t\d+|msc|isc|elsc|el_self", tok), tok
    return tok


def main():
    text = open(SRC).read()
    body = text[text.index("\nscore_brute_matl_cm("):]
    body = body[:body.index("\n}\n")]
    paths = {"J": {}, "L": {}, "R": {}}
    seen = set()
    for m in re.finditer(r"^\s*([JLR])p\[(\d+)\]\s*=\s*([^;]*);", body, flags=re.M):
        mode, idx, expr = m.group(1), int(m.group(2)), m.group(3)
        assert (mode, idx) not in seen, (mode, idx)
        seen.add((mode, idx))
        fs = [factor(t) for t in expr.split("*")]
        n = sum(f in ("msc", "isc", "el_self") for f in fs)
        paths[mode].setdefault(str(n), []).append(sorted(fs))
    out = {"source": "src/itest_brute.c score_brute_matl_cm (Joint, Left and Right marginal paths)",
           "model": ["ROOT", "MATL", "MATL", "END"],
           "paths": paths}
    with open(OUT, "w") as f:   # one path per line
        f.write(json.dumps(out, indent=1).replace("\n     ", " ").replace("\n    ]", " ]"))
    print(OUT, {m: {L: len(v) for L, v in d.items()} for m, d in paths.items()})


if __name__ == "__main__":
    main()
