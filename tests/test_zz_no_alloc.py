"""infx_FastCYKScanNoAlloc, the FastCYKScan variant that allocates no device
memory (its buffers come from the library's fixed scratch array), against
FastCYKScan through the C ABI: best score, hit list, envelope and per-state
maxima bit-identical, on two calls in a row, the second on a shorter target
(the array keeps the first call's contents).  A model and target too large
for the array are refused with eslERANGE before any device work."""
import ctypes as C

import numpy as np
import pytest

import infernal_amd as ia
from helpers import load_cm, random_dsq, hits_equal, describe_hit_diff

pytestmark = pytest.mark.gpu


def noalloc(cm, qdb, dsq, i0, j0, cutoff, env_cutoff=3.0):
    lib = ia.lib()
    f = lib.infx_FastCYKScanNoAlloc
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_float, C.c_void_p,
                  C.c_int, C.c_float, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.POINTER(C.c_float)),
                  C.POINTER(C.c_float)]
    cm.c.search_opts &= ~(ia.CM_SEARCH_INSIDE | ia.CM_SEARCH_CMNOTGREEDY)
    eb = C.create_string_buffer(128)
    th = lib.cm_tophits_Create()
    ei, ej, sc = C.c_int64(), C.c_int64(), C.c_float()
    vsc = C.POINTER(C.c_float)()
    st = f(C.cast(cm.ptr, C.c_void_p), eb, C.cast(cm.c.smx, C.c_void_p), qdb, dsq.ctypes.data, i0, j0, cutoff, th, 0,
           env_cutoff, C.byref(ei), C.byref(ej), C.byref(vsc), C.byref(sc))
    hits = ia._hits_list(th) if st == 0 else None
    lib.cm_tophits_Destroy(th)
    v = None
    if st == 0:
        v = np.array([vsc[k] for k in range(cm.M)], dtype=np.float32)
        C.CDLL(None).free(C.cast(vsc, C.c_void_p))
    return st, eb.value.decode(), {"hits": hits, "envi": ei.value, "envj": ej.value, "sc": sc.value, "vsc": v}


@pytest.mark.parametrize("name,L1,L2", [("tRNA.c.cm", 60000, 7000), ("synth_multiloop350.cm", 200000, 20000)])
def test_no_alloc_equals_fastcykscan(name, L1, L2):
    cm = load_cm(name)
    for k, L in enumerate((L1, L2)):
        dsq = random_dsq(L, seed=900 + k, gc=0.45)
        print("case no-alloc", name, L, flush=True)
        st, msg, g = noalloc(cm, 1, dsq, 1, L, 0.0)
        assert st == 0, msg
        o = ia.FastCYKScan(cm, 1, dsq, 1, L, 0.0, env_cutoff=3.0, want_vsc=True)
        assert np.float32(g["sc"]).view(np.uint32) == np.float32(o["sc"]).view(np.uint32), (g["sc"], o["sc"])
        assert hits_equal(g["hits"], o["hits"]), describe_hit_diff(g["hits"], o["hits"])
        assert (g["envi"], g["envj"]) == (o["envi"], o["envj"])
        assert np.array_equal(g["vsc"].view(np.uint32), o["vsc"].view(np.uint32))


def test_no_alloc_refuses_oversized_targets():
    cm = load_cm("tRNA.c.cm")
    L = 270_000_000
    dsq = np.zeros(L + 2, dtype=np.uint8)
    dsq[0] = dsq[-1] = 255
    st, msg, _ = noalloc(cm, 1, dsq, 1, L, 0.0)
    assert st == 16, (st, msg)   # eslERANGE
    assert "no-allocation scratch" in msg
