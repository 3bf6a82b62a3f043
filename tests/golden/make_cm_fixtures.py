"""Copy the model section (header .. first '//') of shipped Infernal CM files
into tests/golden/cm/.  These are data files of the reference's own test
suite and tutorial; the filter HMM that follows each model is not needed by
the scanners and is dropped.  Run from the repository root:

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


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
