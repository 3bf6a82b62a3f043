"""Copy the model section (header .. first '//') of shipped Infernal CM files
into tests/golden/cm/.  These are data files of the reference's own test
suite and tutorial; the filter HMM that follows each model is not needed by
the scanners and is dropped.  Also extracts the hit table (rank, E-value,
score, start, end, strand) and the searched-residue count of the tutorial's
cmsearch output tutorial/tRNA5-mrum.out into tests/golden/tRNA5_mrum_hits.tsv,
and the stretches of tutorial/mrum-genome.fa around those hits into
tests/golden/mrum_hit_regions.fa.
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


def extract_regions(ref, pad=300):
    """The stretches of tutorial/mrum-genome.fa around the tutorial's hits
    (each hit +-pad, overlapping stretches merged), as FASTA records named
    by their 1-based genome coordinates."""
    here = os.path.dirname(os.path.abspath(__file__))
    seq = []
    with open(os.path.join(ref, "tutorial", "mrum-genome.fa")) as fh:
        for line in fh:
            if not line.startswith(">"):
                seq.append(line.strip())
    genome = "".join(seq)
    spans = []
    with open(os.path.join(here, "tRNA5_mrum_hits.tsv")) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            f = line.split("\t")
            a, b = sorted((int(f[3]), int(f[4])))
            spans.append((max(1, a - pad), min(len(genome), b + pad)))
    spans.sort()
    merged = []
    for a, b in spans:
        if merged and a <= merged[-1][1]:
            merged[-1] = (merged[-1][0], max(merged[-1][1], b))
        else:
            merged.append((a, b))
    with open(os.path.join(here, "mrum_hit_regions.fa"), "w") as fh:
        for a, b in merged:
            fh.write(">mrum_%d_%d\n" % (a, b))
            s = genome[a - 1:b]
            for k in range(0, len(s), 80):
                fh.write(s[k:k + 80] + "\n")
    print("mrum-genome.fa", len(genome), "nt;", len(merged), "regions,", sum(b - a + 1 for a, b in merged), "nt")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    extract_hits(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    extract_regions(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
