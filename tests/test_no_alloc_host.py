"""Host-side checks of infx_FastCYKScanNoAlloc (no GPU needed): a model and
target whose buffers cannot fit the fixed scratch array are refused with
eslERANGE and a message naming the sizes, before any device work; a call
that fits reaches the device check (eslENODEVICE without a GPU)."""
import ctypes as C

import numpy as np
import pytest

import infernal_amd as ia
from helpers import load_cm, random_dsq


def _call(cm, dsq, L):
    lib = ia.lib()
    f = lib.infx_FastCYKScanNoAlloc
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_float, C.c_void_p,
                  C.c_int, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
    cm.c.search_opts &= ~ia.CM_SEARCH_INSIDE
    eb = C.create_string_buffer(128)
    sc = C.c_float()
    st = f(C.cast(cm.ptr, C.c_void_p), eb, C.cast(cm.c.smx, C.c_void_p), 1, dsq.ctypes.data, 1, L, 0.0, None, 0, 0.0,
           None, None, None, C.byref(sc))
    return st, eb.value.decode()


def test_oversized_target_refused():
    if ia.gpu_available():
        pytest.skip("checks the host-side refusal; the GPU run covers it in the no-allocation GPU test")
    cm = load_cm("tRNA.c.cm")
    L = 270_000_000
    dsq = np.zeros(L + 2, dtype=np.uint8)
    dsq[0] = dsq[-1] = 255
    st, msg = _call(cm, dsq, L)
    assert st == 16 and "no-allocation scratch holds 268435456" in msg, (st, msg)


def test_fitting_target_reaches_device_check():
    if ia.gpu_available():
        pytest.skip("no device check to reach on a GPU host")
    cm = load_cm("tRNA.c.cm")
    dsq = random_dsq(5000, seed=1)
    st, msg = _call(cm, dsq, 5000)
    assert st == ia.eslENODEVICE, (st, msg)
