"""E-value calibration and E-values for the scanners (what cmcalibrate and
the search pipeline do around FastCYKScan / FastIInsideScan).

  ExpInfo                 ExpInfo_t (src/infernal.h:499-510) with SetExpInfo,
                          UpdateExpsForDBSize and the survival function
                          (src/stats.c:459-548, esl_exp_surv)
  fit_exp_tail            fit_histogram (src/cmcalibrate.c:1930-2020): the top
                          tailp fraction of the hit scores (Easel's
                          GetTailByMass, count rounded down), maximum
                          likelihood exponential: mu = their minimum,
                          lambda = 1 / (mean - mu)
  tail_mass               tailp = (--ltailn 750 | --gtailn 250) hits per Mb
                          searched / hits collected (src/cmcalibrate.c:1985-1998)
  sample_genomic          SampleGenomicSequenceFromHMM with the five-state
                          generative HMM of CreateGenomicHMM (src/stats.c)
  calibrate               cmcalibrate's four exponential tails (CYK and Inside,
                          glocal and local) from -L 1.6 Mb of generated
                          sequence in 10 kb pieces, QDB2 bands, NULL3 on,
                          every hit (cutoff -inf), scanned on the GPU

The random stream differs from Easel's, so calibrated values agree with the
shipped ECM lines statistically, not digit for digit.
"""
import math
import time
from dataclasses import dataclass

import numpy as np

EXP_MODES = ("GC", "GI", "LC", "LI")
TAILN = {"L": 750, "G": 250}

# CreateGenomicHMM (src/stats.c): start, transition and emission probabilities
# of the five-state HMM cmcalibrate samples its random sequence from.
_GHMM_START = [0.157377049180328, 0.39344262295082, 0.265573770491803, 0.00327868852459016, 0.180327868852459]
_GHMM_TRANS = [
    [0.999483637183643, 0.000317942006440604, 0.000185401071732768, 2.60394763669618e-07, 1.27593434198113e-05],
    [9.76333640771184e-05, 0.99980020511745, 9.191359010352e-05, 7.94413051888677e-08, 1.01684870641751e-05],
    [1.3223694798182e-07, 0.000155642887774602, 0.999700615549769, 9.15079680034191e-05, 5.21013575048369e-05],
    [0.994252873563218, 0.0014367816091954, 0.0014367816091954, 0.0014367816091954, 0.0014367816091954],
    [8.32138798088677e-06, 2.16356087503056e-05, 6.42411152124459e-05, 1.66427759617735e-07, 0.999905635460297],
]
_GHMM_EMIT = [
    [0.370906566523225, 0.129213995153577, 0.130511270043053, 0.369368168280145],
    [0.305194882571888, 0.194580936415687, 0.192343972160245, 0.307880208852179],
    [0.238484980800698, 0.261262845707113, 0.261810301531792, 0.238441871960397],
    [0.699280575539568, 0.00143884892086331, 0.00143884892086331, 0.297841726618705],
    [0.169064007664923, 0.331718611320207, 0.33045427183482, 0.16876310918005],
]


@dataclass
class ExpInfo:
    lambda_: float
    mu_orig: float
    dbsize: float
    nrandhits: int
    tailp: float
    mu_extrap: float = 0.0
    cur_eff_dbsize: float = 0.0

    @classmethod
    def from_fit(cls, lambda_, mu_orig, dbsize, nrandhits, tailp):
        """SetExpInfo: mu extrapolated from the tail to the whole histogram,
        effective database size = the random hits of the calibration."""
        e = cls(lambda_, mu_orig, dbsize, nrandhits, tailp)
        e.mu_extrap = mu_orig - math.log(1.0 / tailp) / lambda_
        e.cur_eff_dbsize = float(nrandhits)
        return e

    @classmethod
    def from_cm_line(cls, fields):
        """The six numbers of an ECM?? line: lambda mu_extrap mu_orig dbsize nrandhits tailp."""
        lam, mux, muo, db, nr, tp = fields
        return cls(float(lam), float(muo), float(db), int(nr), float(tp), float(mux), float(nr))

    def update_for_dbsize(self, Z):
        """UpdateExpsForDBSize: scale the expected number of random hits to a
        search of Z residues (both strands counted)."""
        self.cur_eff_dbsize = (Z / self.dbsize) * float(self.nrandhits)
        return self

    def pvalue(self, sc):
        """esl_exp_surv(sc, mu_extrap, lambda)."""
        sc = np.asarray(sc, dtype=np.float64)
        return np.where(sc < self.mu_extrap, 1.0, np.exp(-self.lambda_ * (sc - self.mu_extrap)))

    def evalue(self, sc):
        return self.pvalue(sc) * self.cur_eff_dbsize

    def score_for_evalue(self, E):
        """E2ScoreGivenExpInfo (src/stats.c:285-300)."""
        return self.mu_extrap + math.log(E / self.cur_eff_dbsize) / (-self.lambda_)


def exp_lines(infos):
    """ECM lines of a CM file header for {mode: ExpInfo} (cm_file.c's order:
    LC, GC, LI, GI)."""
    out = []
    for mode in ("LC", "GC", "LI", "GI"):
        if mode in infos:
            e = infos[mode]
            out.append("ECM%s    %.5f  %10.5f  %10.5f  %10d  %10d  %.6f" % (mode, e.lambda_, e.mu_extrap, e.mu_orig,
                                                                          int(e.dbsize), e.nrandhits, e.tailp))
    return out


def with_exp_lines(cm_text, infos):
    """The model text with its ECM lines replaced by those of infos."""
    lines = [l for l in cm_text.splitlines() if not l.startswith("ECM")]
    k = next(i for i, l in enumerate(lines) if l.split() and l.split()[0] == "CM")
    return "\n".join(lines[:k] + exp_lines(infos) + lines[k:]) + "\n"


def read_exp_infos(cm_path):
    """ECMLC/ECMGC/ECMLI/ECMGI lines of a calibrated CM file, by mode name."""
    out = {}
    with open(cm_path) as fh:
        for line in fh:
            if line.startswith("CM") and line.split()[0] == "CM":
                break
            f = line.split()
            if f and f[0] in ("ECMLC", "ECMGC", "ECMLI", "ECMGI"):
                out[f[0][3:]] = ExpInfo.from_cm_line(f[1:7])
    return out


def tail_mass(mode, total_residues, nhits):
    """Fraction of the score histogram fit: the top (750 local | 250 glocal)
    hits per Mb searched."""
    n = TAILN[mode[0]] * (total_residues / 1e6)
    tp = float(np.float32(np.float32(n) / np.float32(nhits)))
    if tp > 1.0:
        raise ValueError("only %d hits for a %d-hit tail" % (nhits, n))
    return tp


def fit_exp_tail(scores, tailp):
    """ML exponential fit to the top tailp fraction of scores.
    Returns (mu, lambda, n_fit)."""
    x = np.sort(np.asarray(scores, dtype=np.float64))
    n = int(len(x) * tailp)
    if n <= 1:
        raise ValueError("too few points in the tail")
    t = x[len(x) - n:]
    mu = float(t.min())
    lam = 1.0 / float(np.mean(t - mu))
    return mu, lam, n


def sample_genomic(nseq, L, rng):
    """nseq sequences of L residues (codes 0..3) from the generative HMM,
    sampled together: state path, then one emission per position."""
    start = np.asarray(_GHMM_START) / np.sum(_GHMM_START)
    trans = np.asarray(_GHMM_TRANS)
    trans = trans / trans.sum(1, keepdims=True)
    emit = np.asarray(_GHMM_EMIT)
    emit = emit / emit.sum(1, keepdims=True)
    tcdf = np.cumsum(trans, 1)
    ecdf = np.cumsum(emit, 1)
    s = (rng.random(nseq)[:, None] > np.cumsum(start)[None, :]).sum(1).clip(0, 4)
    out = np.empty((nseq, L), dtype=np.uint8)
    u_e = rng.random((L, nseq))
    u_t = rng.random((L, nseq))
    for x in range(L):
        out[:, x] = (u_e[x][:, None] > ecdf[s]).sum(1).clip(0, 3)
        s = (u_t[x][:, None] > tcdf[s]).sum(1).clip(0, 4)
    return out


def calibrate(cm_path, L_mb=1.6, seed=181, modes=("GC", "GI", "LC", "LI"), chunk=10000, null3=True, joined=False,
              progress=None):
    """cmcalibrate's exponential tails for the model in cm_path, the scans
    on the GPU.  Returns {mode: ExpInfo} and the hit scores per mode.
    joined: scan the generated pieces end to end as one target (one launch
    per mode instead of one per piece; hits may then span the junctions
    between pieces, a few in a thousand for W of a few hundred).
    null3: score hits with the NULL3 correction, as cmcalibrate does.  The
    scans report every hit (cutoff -inf, src/cmcalibrate.c:1845); with NULL3
    every banded root cell goes to the host replay (about L * W of them),
    without it the scanners' greedy record filter keeps a few per row.
    progress(mode, seconds, hits) is called after each mode."""
    from . import CM, FastCYKScan, FastIInsideScan, dsq_from_codes
    nseq = int(L_mb * 1e6 / chunk + 0.5)
    total = nseq * chunk
    rng = np.random.default_rng(seed)
    seqs = sample_genomic(nseq, chunk, rng)
    if joined:
        seqs = [seqs.reshape(-1)]
        chunk = total
    dsqs = [dsq_from_codes(s) for s in seqs]
    res, all_scores = {}, {}
    for mode in modes:
        t0 = time.perf_counter()
        cm = CM.read(cm_path)
        cm.configure(local=(mode[0] == "L"))
        scan = FastIInsideScan if mode[1] == "I" else FastCYKScan
        scores = []
        for d in dsqs:
            r = scan(cm, 2, d, 1, chunk, -math.inf, do_null3=null3, want_hits=True)
            scores.extend(h[4] for h in r["hits"])
        tp = tail_mass(mode, total, len(scores))
        mu, lam, _ = fit_exp_tail(scores, tp)
        res[mode] = ExpInfo.from_fit(lam, mu, float(total), len(scores), tp)
        all_scores[mode] = np.asarray(scores)
        if progress:
            progress(mode, time.perf_counter() - t0, len(scores))
    return res, all_scores
