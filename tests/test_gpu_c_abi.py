"""The five scanners called from C (tests/c/abi_check.c, gcc, only
include/infernal_gpu.h) on the GPU, checked against the oracle: best scores
bit-exact, best mode, envelope and hit lists."""
import os
import subprocess

import numpy as np
import pytest

import infernal_amd as ia
import oracle_bind as ob
from helpers import load_cm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CDIR = os.path.join(ROOT, "tests", "c")


def _parse(out):
    res, cur = {}, None
    for ln in out.splitlines():
        f = ln.split()
        if ln.startswith("  hit"):
            cur["hits"].append((int(f[1]), int(f[2]), int(f[3]), int(f[4]), int(f[5], 16)))
        elif not ln.startswith(" "):
            cur = {"status": int(f[2]), "bits": int(f[6], 16), "mode": int(f[8]), "env": (int(f[10]), int(f[11])), "hits": []}
            res[f[0]] = cur
    return res


def test_c_caller_matches_oracle(tmp_path):
    exe = os.path.join(CDIR, "abi_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", CDIR, "abi_check"], check=True, capture_output=True)
    L = 600
    rng = np.random.default_rng(11)
    codes = rng.integers(0, 4, L).astype(np.uint8)
    seq = tmp_path / "seq.txt"
    seq.write_text("".join("ACGU"[x] for x in codes) + "\n")
    print("case abi_check tRNA.c.cm L=%d" % L, flush=True)
    r = subprocess.run([exe, "scan", os.path.join(ROOT, "tests", "golden", "cm", "tRNA.c.cm"), str(seq)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _parse(r.stdout)
    cm = load_cm("tRNA.c.cm", local=True, trunc=True)
    dsq = ia.dsq_from_codes(codes)
    cases = {"FastCYKScan": (ob.ORACLE_CYK, 1, 1), "FastFInsideScan": (ob.ORACLE_FINSIDE, 2, 1),
             "FastIInsideScan": (ob.ORACLE_IINSIDE, 2, 1),
             "RefTrCYKScan": (ob.ORACLE_TRCYK, 1, ia.PLI_PASS_5P_AND_3P_FORCE),
             "RefITrInsideScan": (ob.ORACLE_TRIINSIDE, 2, ia.PLI_PASS_5P_ONLY_FORCE)}
    for name, (algo, q, p) in cases.items():
        o = ob.scan(cm, algo, q, dsq, 1, L, 0.0, env_cutoff=2.0, pass_idx=p)
        g = got[name]
        assert g["status"] == 0, (name, r.stdout)
        assert g["bits"] == int(np.float32(o["sc"]).view(np.uint32)), (name, g, o["sc"])
        assert g["env"] == (o["envi"], o["envj"]), name
        if algo in (ob.ORACLE_TRCYK, ob.ORACLE_TRIINSIDE):
            assert g["mode"] == o["mode"], name
        oh = [(h[0], h[1], h[2], h[3], int(np.float32(h[4]).view(np.uint32))) for h in o["hits"]]
        assert g["hits"] == oh, (name, g["hits"], oh)


def test_dlopen_binding_matches_oracle():
    """tests/c/dlopen_check.c: a host that defines the reference's own names
    (FastCYKScan, FreeCM, cm_Configure, ...) loads the library with
    dlopen(RTLD_LOCAL) and calls cm_Configure, FastCYKScan and RefTrCYKScan
    through dlsym on a tRNA; both best scores bit-exact with the oracle."""
    exe = os.path.join(CDIR, "dlopen_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", CDIR, "dlopen_check"], check=True, capture_output=True)
    lib = os.path.join(ROOT, "infernal_amd", "csrc", "libinfernal_gpu.so")
    print("case dlopen_check tRNA.c.cm", flush=True)
    r = subprocess.run([exe, lib, os.path.join(ROOT, "tests", "golden", "cm", "tRNA.c.cm")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "dlopen ok" in r.stdout, r.stdout + r.stderr
    got = {ln.split()[0]: ln.split() for ln in r.stdout.splitlines() if "status" in ln}
    seq = "GCCGGGAUAGCUCAGUUGGUAGAGCAGAGGACUGAAAAUCCUCGUGUCACCGGUUCGAUUCCGGUCCCGGCACCA"
    dsq = ia.digitize(seq)
    L = len(seq)
    cm = load_cm("tRNA.c.cm", local=True, trunc=True)
    o1 = ob.scan(cm, ob.ORACLE_CYK, 1, dsq, 1, L, 0.0)
    o2 = ob.scan(cm, ob.ORACLE_TRCYK, 1, dsq, 1, L, 0.0, pass_idx=ia.PLI_PASS_5P_AND_3P_FORCE)
    assert int(got["FastCYKScan"][2]) == 0 and int(got["RefTrCYKScan"][2]) == 0, r.stdout
    assert int(got["FastCYKScan"][4], 16) == int(np.float32(o1["sc"]).view(np.uint32)), (r.stdout, o1["sc"])
    assert int(got["RefTrCYKScan"][4], 16) == int(np.float32(o2["sc"]).view(np.uint32)), (r.stdout, o2["sc"])
    assert int(got["RefTrCYKScan"][6]) == o2["mode"], r.stdout
    assert o1["sc"] > 20.0
