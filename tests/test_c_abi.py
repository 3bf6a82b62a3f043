"""The C ABI from a C caller: tests/c/abi_check.c is compiled with gcc
against include/infernal_gpu.h alone and linked to libinfernal_gpu.so, as
Infernal's cm_pipeline.c would be.

* Without a GPU every scanner returns eslENODEVICE through its declared
  prototype (argument checking and profile building still run).
* The truncated scan matrix's teardown, under AddressSanitizer: the host code
  is rebuilt with g++ -fsanitize=address (host objects only; the gfx950
  kernel objects are the library's), truncated-scanner resources are attached
  to a configured model as a GPU scan leaves them, and FreeCM must release
  them without a heap error.  Before the fix, cm_tr_scan_mx_Destroy deleted
  those resources as the standard matrix's type, and this run reported
  new-delete-type-mismatch.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CDIR = os.path.join(ROOT, "tests", "c")
CMDIR = os.path.join(ROOT, "tests", "golden", "cm")


def build(target):
    path = os.path.join(CDIR, target)
    if not os.path.exists(path):
        subprocess.run(["make", "-C", CDIR, target], check=True, capture_output=True)
    return path


def test_c_caller_builds_and_calls_every_scanner(tmp_path):
    exe = build("abi_check")
    seq = tmp_path / "seq.txt"
    rng = np.random.default_rng(3)
    seq.write_text("".join("ACGU"[x] for x in rng.integers(0, 4, 400)) + "\n")
    r = subprocess.run([exe, "scan", os.path.join(CMDIR, "tRNA.c.cm"), str(seq)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    names = [ln.split()[0] for ln in r.stdout.splitlines() if not ln.startswith(" ")]
    assert names == ["FastCYKScan", "FastFInsideScan", "FastIInsideScan", "RefTrCYKScan", "RefITrInsideScan"]
    status = {int(ln.split()[2]) for ln in r.stdout.splitlines() if not ln.startswith(" ")}
    assert status in ({0}, {100}), r.stdout


def test_truncated_matrix_teardown_under_asan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "infernal_amd", "csrc")], check=True, capture_output=True)
    exe = os.path.join(CDIR, "abi_check_asan")
    subprocess.run(["make", "-C", CDIR, "abi_check_asan"], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1")
    for name in ("tRNA.c.cm", "synth_nested1500.cm"):
        r = subprocess.run([exe, "teardown", os.path.join(CMDIR, name)], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0 and "teardown ok" in r.stdout, r.stdout + r.stderr[-3000:]
        assert "AddressSanitizer" not in r.stderr, r.stderr[-3000:]


_REPLAY_DRIVER = r"""
import ctypes as C, sys
import numpy as np
lib = C.CDLL(sys.argv[1])
class CAND(C.Structure):
    _fields_ = [("j", C.c_int32), ("d", C.c_int32), ("sc", C.c_float), ("r", C.c_int32), ("mode", C.c_int32)]
lib.cm_tophits_Create.restype = C.c_void_p
lib.cm_tophits_Destroy.argtypes = [C.c_void_p]
f = lib.infx_resolve_hits
f.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
              C.c_int, C.c_float, C.c_int, C.c_int, C.c_void_p]
rng = np.random.default_rng(5)
i0, j0, W = 1000001, 1400000, 300
js = rng.integers(i0, j0 + 1, 400000)
ds = np.minimum(rng.integers(1, W + 1, len(js)), js - i0 + 1)
_, first = np.unique(js.astype(np.int64) * 1024 + ds, return_index=True)
js, ds = js[first], ds[first]
for greedy in (1, 0):
    arr = (CAND * len(js))()
    a = np.frombuffer(arr, dtype=[("j", "<i4"), ("d", "<i4"), ("sc", "<f4"), ("r", "<i4"), ("mode", "<i4")])
    a["j"], a["d"], a["sc"], a["r"], a["mode"] = js, ds, (rng.integers(-40, 80, len(js)) / 2.0), 0, 3
    th = lib.cm_tophits_Create()
    eb = C.create_string_buffer(128)
    st = f(None, eb, 0, 0, C.cast(arr, C.c_void_p), len(js), None, i0, j0, W, 0.0, 0, greedy, th)
    assert st == 0, eb.value
    lib.cm_tophits_Destroy(th)
print("replay ok", len(js))
"""


def test_hit_replay_under_asan():
    """The host replay (radix orderings, compact hit records, bitmap pass,
    gamma path) on a set above the radix threshold, in the library's host
    code rebuilt with AddressSanitizer and loaded into a Python process with
    the ASan runtime preloaded."""
    subprocess.run(["make", "-C", os.path.join(ROOT, "infernal_amd", "csrc")], check=True, capture_output=True)
    subprocess.run(["make", "-C", CDIR, "asan_lib"], check=True, capture_output=True)
    lib = os.path.join(CDIR, "_asan", "libinfernal_gpu_asan.so")
    asan_rt = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", LD_PRELOAD=asan_rt)
    r = subprocess.run([sys.executable, "-c", _REPLAY_DRIVER, lib], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "replay ok" in r.stdout, r.stdout + r.stderr[-3000:]
    assert "AddressSanitizer" not in r.stderr, r.stderr[-3000:]


def test_dlopen_binding_beside_same_named_host_functions():
    """tests/c/dlopen_check.c defines FastCYKScan, FreeCM, cm_Configure,
    cm_scan_mx_Create, cm_tophits_* ... itself (as an Infernal binary does) and
    exports them (-rdynamic); it loads the library with dlopen(RTLD_LOCAL) and
    calls cm_Configure, FastCYKScan and RefTrCYKScan through dlsym.  Its own
    definitions abort if called: the library's calls among its own functions
    must bind inside it (the library is linked -Bsymbolic)."""
    exe = build("dlopen_check")
    lib = os.path.join(ROOT, "infernal_amd", "csrc", "libinfernal_gpu.so")
    r = subprocess.run([exe, lib, os.path.join(CMDIR, "tRNA.c.cm")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "dlopen ok" in r.stdout, r.stdout + r.stderr
    status = {int(ln.split()[2]) for ln in r.stdout.splitlines() if "status" in ln}
    assert status in ({0}, {100}), r.stdout
