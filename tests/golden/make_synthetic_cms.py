"""Build the three synthetic benchmark models with infernal_amd.cmbuild and
write them to tests/golden/cm/.  The Dirichlet-mixture priors are read from
the reference's default prior file (src/default.prior) at generation time;
only the generated models are committed.  Run from the repository root:

    python tests/golden/make_synthetic_cms.py /root/reference

Models (structure, alignment):
  synth_cloverleaf76   76-nt cloverleaf, stems 7/4/5/5 bp, 50 sequences, 25% substitution
  synth_multiloop350   350-nt consensus, 3-way multiloop, 12 stems of 4-10 bp, 50 sequences, 25%
  synth_nested1500     1500-nt consensus, 45 nested stems of 3-12 bp, 200 sequences, 20%
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from infernal_amd import cmbuild as cb  # noqa: E402

MODELS = [
    ("synth_cloverleaf76", 1, lambda rng: cb.cloverleaf_76(), 50, 0.25),
    ("synth_multiloop350", 2, lambda rng: cb.multiloop_structure(350, 12, rng, bp_range=(4, 10)), 50, 0.25),
    ("synth_nested1500", 3, lambda rng: cb.nested_structure(1500, 45, rng, bp_range=(3, 12)), 200, 0.20),
]


def main(ref):
    prior = os.path.join(ref, "src", "default.prior")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cm")
    for name, seed, make_ss, nseq, mut in MODELS:
        t = time.time()
        rng = np.random.default_rng(seed)
        ss = make_ss(rng)
        aln = cb.synthetic_alignment(ss, nseq, mut, rng)
        text, info = cb.build_cm(ss, aln, name, prior)
        with open(os.path.join(out, name + ".cm"), "w") as fh:
            fh.write(text)
        print(name, info, "%.1f s" % (time.time() - t))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
