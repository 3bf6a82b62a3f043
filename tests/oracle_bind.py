"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: the oracle is the CPU restatement the GPU scanners are
checked against; only tests, smoke() and bench.py's CPU baseline use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

ORACLE_CYK, ORACLE_FINSIDE, ORACLE_IINSIDE, ORACLE_TRCYK, ORACLE_TRIINSIDE = range(5)


class ORACLE_HIT(C.Structure):
    _fields_ = [("start", C.c_int64), ("stop", C.c_int64), ("root", C.c_int), ("mode", C.c_int), ("score", C.c_float)]


class ORACLE_RESULT(C.Structure):
    _fields_ = [("nhits", C.c_int), ("hits", C.POINTER(ORACLE_HIT)), ("envi", C.c_int64), ("envj", C.c_int64),
                ("sc", C.c_float), ("mode", C.c_int), ("vsc", C.POINTER(C.c_float)), ("W", C.c_int),
                ("rows_sc", C.POINTER(C.c_float)), ("rows_r", C.POINTER(C.c_int)), ("rows_mode", C.POINTER(C.c_char)),
                ("rows_msc", C.POINTER(C.c_float))]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
        _lib = C.CDLL(ORACLE_SO)
        _lib.oracle_scan.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_int64,
                                     C.c_float, C.c_int, C.c_int, C.c_float, C.c_int, C.POINTER(ORACLE_RESULT),
                                     C.c_char_p]
        _lib.oracle_result_free.argtypes = [C.POINTER(ORACLE_RESULT)]
        _lib.oracle_flogsum_table.restype = C.POINTER(C.c_float)
        _lib.oracle_ilogsum_table.restype = C.POINTER(C.c_int)
    return _lib


def scan(cm, algo, qdbidx, dsq, i0, j0, cutoff, do_null3=False, greedy=True, env_cutoff=0.0, want_rows=False,
         pass_idx=1):
    lib = load()
    res = ORACLE_RESULT()
    eb = C.create_string_buffer(128)
    dsq = np.ascontiguousarray(dsq, dtype=np.uint8)
    st = lib.oracle_scan(C.cast(cm.ptr, C.c_void_p), algo, qdbidx, pass_idx, dsq.ctypes.data, i0, j0, cutoff,
                         int(do_null3), int(greedy), env_cutoff, int(want_rows), C.byref(res), eb)
    if st != 0:
        raise RuntimeError("oracle status %d: %s" % (st, eb.value.decode()))
    M = cm.M
    out = {
        "hits": [(res.hits[h].start, res.hits[h].stop, res.hits[h].root, res.hits[h].mode, res.hits[h].score)
                 for h in range(res.nhits)],
        "envi": res.envi, "envj": res.envj, "sc": res.sc, "mode": res.mode,
        "vsc": np.array([res.vsc[v] for v in range(M)], dtype=np.float32),
    }
    if want_rows:
        n = (j0 - i0 + 1) * (res.W + 1)
        out["W"] = res.W
        out["rows_sc"] = np.ctypeslib.as_array(res.rows_sc, shape=(n,)).copy().reshape(j0 - i0 + 1, res.W + 1)
        out["rows_r"] = np.ctypeslib.as_array(res.rows_r, shape=(n,)).copy().reshape(j0 - i0 + 1, res.W + 1)
        if res.rows_mode:
            out["rows_mode"] = np.frombuffer(C.string_at(res.rows_mode, n), dtype=np.uint8).copy().reshape(j0 - i0 + 1, res.W + 1)
        if res.rows_msc:
            out["rows_msc"] = np.ctypeslib.as_array(res.rows_msc, shape=(n * 4,)).copy().reshape(j0 - i0 + 1, res.W + 1, 4)
    lib.oracle_result_free(C.byref(res))
    return out
