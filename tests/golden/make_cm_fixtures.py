"""Copy the model section (header .. first '//') of shipped Infernal CM files
into tests/golden/cm/.  These are data files of the reference's own test
suite and tutorial; the filter HMM that follows each model is not needed by
the scanners and is dropped.  Also extracts the hit table (rank, E-value,
score, start, end, strand) and the searched-residue count of the tutorial's
cmsearch output tutorial/tRNA5-mrum.out into tests/golden/tRNA5_mrum_hits.tsv.
Run from the repository root:

    python tests/golden/make_cm_fixtures.py /root/reference
"""
import os
import sys

FILES = [
    "testsuite/tRNA.c.cm",
    "testsuite/Vault.c.cm",
    "tutorial/5S_rRNA.c.cm",
    "tutorial/Cobalamin.c.cm",
    "testsuite/Plant_SRP.c.cm",
    "testsuite/rnaseP-eubact.c.cm",
    "tutorial/tRNA5.c.cm",
]


def main(ref):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cm")
    os.makedirs(out, exist_ok=True)
    for rel in FILES:
        lines = []
        with open(os.path.join(ref, rel)) as fh:
            for line in fh:
                lines.append(line)
                if line.strip() == "//":
                    break
        with open(os.path.join(out, os.path.basename(rel)), "w") as fh:
            fh.writelines(lines)
        print(rel, len(lines), "lines")


def extract_hits(ref):
    import re
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tRNA5_mrum_hits.tsv")
    rows, z = [], None
    with open(os.path.join(ref, "tutorial", "tRNA5-mrum.out")) as fh:
        for line in fh:
            m = re.match(r"\s*\((\d+)\)\s+[!?]\s+(\S+)\s+(\S+)\s+\S+\s+\S+\s+(\d+)\s+(\d+)\s+([+-])", line)
            if m:
                rows.append(m.groups())
            m = re.search(r"Target sequences:\s+\d+\s+\((\d+) residues searched\)", line)
            if m and z is None:
                z = m.group(1)
    with open(out, "w") as fh:
        fh.write("# cmsearch tRNA5.c.cm mrum-genome.fa (tutorial/tRNA5-mrum.out); residues searched %s\n" % z)
        fh.write("#rank\tevalue\tscore\tstart\tend\tstrand\n")
        for r in rows:
            fh.write("\t".join(r) + "\n")
    print("tRNA5-mrum.out", len(rows), "hits")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    extract_hits(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
