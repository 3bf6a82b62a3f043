"""Build the many-models workload's synthetic covariance models with
infernal_amd.cmbuild: <n> models whose consensus lengths are log-uniform in
[lmin, lmax] nt, each a random nested stem tree of 1 to min(60, clen/25)
stems of 3-10 bp, built from a seeded 20-sequence alignment at 20%
substitution (compensatory pair changes) with the default Dirichlet priors,
which are read from the reference's src/default.prior.  The models are
written gzip-compressed to <outdir> (they are large: ~30 MB for 200; they are
inputs for a benchmark leg, not fixtures).  Run from the repository root:

    python tests/golden/make_batch_cms.py /root/reference <outdir> [n] [seed]
"""
import gzip
import multiprocessing as mp
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from infernal_amd import cmbuild as cb  # noqa: E402

LMIN, LMAX = 50, 1500


def plan(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        clen = int(round(np.exp(rng.uniform(np.log(LMIN), np.log(LMAX)))))
        nstems = int(rng.integers(1, max(1, min(60, clen // 25)) + 1))
        out.append((i, clen, nstems, int(rng.integers(1 << 30))))
    return out


def build(job):
    i, clen, nstems, s, prior, outdir = job
    sys.setrecursionlimit(100000)      # guide trees of deeply nested stems recurse per node
    name = "batch%03d_%d" % (i, clen)
    if os.path.exists(os.path.join(outdir, name + ".cm.gz")):
        return name, clen, 0, 0, nstems
    rng = np.random.default_rng(s)
    while True:                          # fewer stems until the structure fits the length
        try:
            ss = cb.nested_structure(clen, nstems, rng, bp_range=(3, 10))
            break
        except ValueError:
            if nstems == 1:
                raise
            nstems -= 1
    aln = cb.synthetic_alignment(ss, 20, 0.20, rng)
    text, info = cb.build_cm(ss, aln, name, prior)
    with gzip.open(os.path.join(outdir, name + ".cm.gz"), "wt") as fh:
        fh.write(text)
    return name, info["clen"], info["M"], info["W"], nstems


def main(ref, outdir, n=200, seed=2026):
    prior = os.path.join(ref, "src", "default.prior")
    os.makedirs(outdir, exist_ok=True)
    jobs = [(i, clen, ns, s, prior, outdir) for (i, clen, ns, s) in plan(n, seed)]
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        for name, clen, M, W, ns in pool.imap_unordered(build, jobs):
            print(name, "clen", clen, "M", M, "W", W, "stems", ns, flush=True)


if __name__ == "__main__":
    a = sys.argv
    main(a[1], a[2], int(a[3]) if len(a) > 3 else 200, int(a[4]) if len(a) > 4 else 2026)
