"""infx_FastScanBatch: several models' windows of one target in one launch.

Each model's hits and best score must equal, bit for bit, those of
infx_FastScanWindows over that model's windows alone (itself checked against
the oracle in test_gpu_parity_std.py).  The models differ in size (tRNA,
the 350-nt multiloop, Cobalamin with bifurcations); windows of different
models overlap; a forced short chunk length gives many chunks per window.
"""
import os

import numpy as np
import pytest

import infernal_amd as ia
from helpers import CMDIR, load_cm, hits_equal, describe_hit_diff

pytestmark = pytest.mark.gpu


def _target(L, seed):
    from infernal_amd import synth
    frags = []
    for k, name in enumerate(("tRNA.c.cm", "synth_multiloop350.cm")):
        frags += synth.homologs(ia.CM.read(os.path.join(CMDIR, name)), 3, seed=seed + k)
    bg = synth.markov2_background(L, gc=0.45, seed=seed + 7)
    codes, _ = synth.plant(bg, frags, seed=seed + 9)
    return ia.dsq_from_codes(codes)


WINDOWS = {0: [(1, 5000), (6000, 9000), (15000, 15100)],
           1: [(2000, 12000), (12500, 14000)],
           2: [(500, 8000), (10000, 20000)]}


@pytest.mark.parametrize("algo", ["cyk", "finside", "iinside"])
def test_batch_equals_per_model_windows(algo):
    names = ["tRNA.c.cm", "synth_multiloop350.cm", "Cobalamin.c.cm"]
    cms = [load_cm(n) for n in names]
    dsq = _target(20000, seed=61)
    qdb = 1 if algo == "cyk" else 2
    cutoffs = [-5.0, 0.0, -2.0]
    wins = [(i, j, m) for m, ws in WINDOWS.items() for (i, j) in ws]
    print("case batch", algo, len(wins), flush=True)
    ia.set_launch_params(chunk_len=900)
    try:
        got = ia.ScanBatch(algo, cms, qdb, dsq, wins, cutoffs)
    finally:
        ia.set_launch_params(0, 0)
    key = lambda h: (h[0], h[1])
    for m, cm in enumerate(cms):
        want = ia.ScanWindows(algo, cm, qdb, dsq, WINDOWS[m], cutoffs[m])
        g = sorted(got[m]["hits"], key=key)
        w = sorted(want["hits"], key=key)
        assert hits_equal(g, w), (names[m], describe_hit_diff(g, w))
        assert np.float32(got[m]["sc"]).view(np.uint32) == np.float32(want["sc"]).view(np.uint32), (names[m], got[m]["sc"], want["sc"])
    assert sum(len(x["hits"]) for x in got) > 0
