"""Shared test helpers: fixtures, sequence generators, comparisons."""
import os

import numpy as np

import infernal_amd as ia

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CMDIR = os.path.join(ROOT, "tests", "golden", "cm")

_cache = {}


def cm_path(name):
    return os.path.join(CMDIR, name)


def load_cm(name, local=True, trunc=False):
    key = (name, local, trunc)
    if key not in _cache:
        cm = ia.CM.read(os.path.join(CMDIR, name))
        cm.configure(local=local, trunc=trunc, scanmx=True, trscanmx=trunc)
        _cache[key] = cm
    return _cache[key]


def random_dsq(L, seed, gc=None, degenerate=0.0):
    rng = np.random.default_rng(seed)
    if gc is None:
        codes = rng.integers(0, 4, L)
    else:
        p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
        codes = rng.choice(4, size=L, p=p)
    if degenerate > 0:
        mask = rng.random(L) < degenerate
        codes[mask] = rng.integers(5, 16, mask.sum())
    return ia.dsq_from_codes(codes.astype(np.uint8))


def hits_equal(a, b):
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        if x[0] != y[0] or x[1] != y[1] or x[2] != y[2] or x[3] != y[3]:
            return False
        if np.float32(x[4]).view(np.uint32) != np.float32(y[4]).view(np.uint32):
            return False
    return True


def describe_hit_diff(a, b, n=5):
    out = ["len %d vs %d" % (len(a), len(b))]
    for k in range(min(max(len(a), len(b)), n)):
        out.append("%r | %r" % (a[k] if k < len(a) else None, b[k] if k < len(b) else None))
    return "\n".join(out)
