"""infernal_amd -- MI355X (gfx950) QDB scanning CYK/Inside for covariance models.

Python view of the C ABI declared in include/infernal_gpu.h (libinfernal_gpu.so,
built in infernal_amd/csrc).  The scanner functions keep the reference's names
and argument meaning (src/cm_dpsearch.c FastCYKScan / FastFInsideScan /
FastIInsideScan, src/cm_dpsearch_trunc.c RefTrCYKScan / RefITrInsideScan).

There is no CPU fallback: importing works without a GPU, but a scan raises
ScanError (eslENODEVICE) when no gfx9 HIP device is present, and importing
raises ImportError when the native library has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "csrc", "libinfernal_gpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libinfernal_gpu.so is not built (%s); run __graft_entry__.build() or make -C infernal_amd/csrc" % LIB_PATH)

_lib = C.CDLL(LIB_PATH)

# ---------------------------------------------------------------------------
# constants (include/infernal_gpu.h)
eslOK, eslEINCOMPAT, eslEINVAL, eslENODEVICE = 0, 10, 11, 100
SMX_NOQDB, SMX_QDB1_TIGHT, SMX_QDB2_LOOSE = 0, 1, 2
CMH_BITS, CMH_LOCAL_BEGIN, CMH_LOCAL_END = 1 << 0, 1 << 10, 1 << 11
CM_CONFIG_LOCAL, CM_CONFIG_TRUNC, CM_CONFIG_SCANMX, CM_CONFIG_TRSCANMX = 1 << 0, 1 << 5, 1 << 6, 1 << 7
CM_SEARCH_INSIDE, CM_SEARCH_CMNOTGREEDY, CM_SEARCH_NULL3 = 1 << 5, 1 << 8, 1 << 11
TRMODE_T, TRMODE_R, TRMODE_L, TRMODE_J, TRMODE_UNKNOWN = 0, 1, 2, 3, 4
PLI_PASS_STD_ANY, PLI_PASS_5P_ONLY_FORCE, PLI_PASS_3P_ONLY_FORCE = 1, 2, 3
PLI_PASS_5P_AND_3P_FORCE, PLI_PASS_5P_AND_3P_ANY = 4, 5
IMPOSSIBLE = -1e36
D_st, MP_st, ML_st, MR_st, IL_st, IR_st, S_st, E_st, B_st = range(9)
ERRBUF = 128
RNA = "ACGU-RYMKSWHBVDN*~"

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_pp = C.c_void_p


class CM_QDBINFO(C.Structure):
    _fields_ = [("dmin1", _ip), ("dmax1", _ip), ("dmin2", _ip), ("dmax2", _ip),
                ("beta1", C.c_double), ("beta2", C.c_double)]


class CMEmitMap(C.Structure):
    _fields_ = [("clen", C.c_int), ("lpos", _ip), ("rpos", _ip), ("epos", _ip)]


class CM_t(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("M", C.c_int), ("nodes", C.c_int), ("clen", C.c_int), ("W", C.c_int),
        ("pbegin", C.c_float), ("pend", C.c_float),
        ("beta_W", C.c_double),
        ("null3_omega", C.c_float),
        ("null", _fp),
        ("sttype", C.POINTER(C.c_char)), ("stid", C.POINTER(C.c_char)),
        ("ndidx", _ip), ("cfirst", _ip), ("cnum", _ip), ("plast", _ip), ("pnum", _ip),
        ("nodemap", _ip), ("ndtype", C.POINTER(C.c_char)),
        ("t", C.POINTER(_fp)), ("e", C.POINTER(_fp)),
        ("begin", _fp), ("end", _fp),
        ("tsc", C.POINTER(_fp)), ("esc", C.POINTER(_fp)),
        ("lmesc", C.POINTER(_fp)), ("rmesc", C.POINTER(_fp)),
        ("beginsc", _fp), ("endsc", _fp),
        ("itsc", C.POINTER(_ip)), ("iesc", C.POINTER(_ip)), ("ilmesc", C.POINTER(_ip)), ("irmesc", C.POINTER(_ip)),
        ("ibeginsc", _ip), ("iendsc", _ip),
        ("oesc", C.POINTER(_fp)), ("ioesc", C.POINTER(_ip)),
        ("el_selfsc", C.c_float), ("iel_selfsc", C.c_int),
        ("root_trans", _fp),
        ("flags", C.c_int), ("config_opts", C.c_int), ("search_opts", C.c_int),
        ("qdbinfo", C.POINTER(CM_QDBINFO)),
        ("emap", C.POINTER(CMEmitMap)),
        ("trp", _pp), ("smx", _pp), ("trsmx", _pp),
    ]


class CM_HIT(C.Structure):
    _fields_ = [("start", C.c_int64), ("stop", C.c_int64), ("srcL", C.c_int64), ("hit_idx", C.c_int64),
                ("cm_idx", C.c_int64), ("seq_idx", C.c_int64), ("pass_idx", C.c_int), ("in_rc", C.c_int),
                ("root", C.c_int), ("mode", C.c_char), ("score", C.c_float), ("hmmonly", C.c_int), ("flags", C.c_int)]


class CM_TOPHITS(C.Structure):
    _fields_ = [("hit", C.POINTER(C.POINTER(CM_HIT))), ("unsrt", C.POINTER(CM_HIT)),
                ("Nalloc", C.c_uint64), ("N", C.c_uint64), ("is_sorted_for_overlap_removal", C.c_int)]


_CMP = C.POINTER(CM_t)
_THP = C.POINTER(CM_TOPHITS)

_lib.infx_cm_Read.argtypes = [C.c_char_p, C.POINTER(_CMP), C.c_char_p]
_lib.infx_cm_ReadText.argtypes = [C.c_char_p, C.POINTER(_CMP), C.c_char_p]
_lib.FreeCM.argtypes = [_CMP]
_lib.cm_Configure.argtypes = [_CMP, C.c_char_p, C.c_int]
_lib.cm_tophits_Create.restype = _THP
_lib.cm_tophits_Destroy.argtypes = [_THP]
_lib.infx_abc_Digitize.argtypes = [C.c_char_p, C.c_int64, C.c_void_p]
_lib.infx_gpu_available.argtypes = [C.c_char_p]
_lib.infx_last_call_timing.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int)]
_lib.infx_set_launch_params.argtypes = [C.c_int, C.c_int]
_lib.infx_CountSearchCells.argtypes = [_CMP, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
_lib.infx_CountSearchWork.argtypes = [_CMP, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
_lib.FLogsum.argtypes = [C.c_float, C.c_float]
_lib.FLogsum.restype = C.c_float
_lib.ILogsum.argtypes = [C.c_int, C.c_int]
_lib.ILogsum.restype = C.c_int
_lib.cm_hit_AllowTruncation.argtypes = [_CMP, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_char, C.c_int]

_STD_ARGS = [_CMP, C.c_char_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_float, _THP, C.c_int,
             C.c_float, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(_fp), C.POINTER(C.c_float)]
_TR_ARGS = [_CMP, C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_float, _THP, C.c_int,
            C.c_float, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(_fp), C.c_char_p, C.POINTER(C.c_float)]
for _n in ("FastCYKScan", "FastFInsideScan", "FastIInsideScan"):
    getattr(_lib, _n).argtypes = _STD_ARGS
for _n in ("RefTrCYKScan", "RefITrInsideScan"):
    getattr(_lib, _n).argtypes = _TR_ARGS


class ScanError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("status %d: %s" % (status, msg))
        self.status = status


def lib():
    """The loaded ctypes library (for tests that call the C ABI directly)."""
    return _lib


def gpu_available():
    eb = C.create_string_buffer(ERRBUF)
    return bool(_lib.infx_gpu_available(eb))


def set_launch_params(chunk_len=0, waves_per_block=0):
    _lib.infx_set_launch_params(int(chunk_len), int(waves_per_block))


def pin_target(dsq):
    """infx_target_Pin: keep a device copy of dsq[1..L] (dsq as from
    dsq_from_codes, sentinels at both ends); scans of this same array read it
    in place.  Keep the array alive and unchanged while pinned."""
    dsq = np.asarray(dsq)
    if dsq.dtype != np.uint8 or not dsq.flags.c_contiguous:
        raise ValueError("pin_target: dsq must be a contiguous uint8 array")
    eb = C.create_string_buffer(ERRBUF)
    st = _lib.infx_target_Pin(C.c_void_p(dsq.ctypes.data), C.c_int64(len(dsq) - 2), eb)
    if st != eslOK:
        raise ScanError(st, eb.value.decode(errors="replace"))


def unpin_target():
    _lib.infx_target_Unpin()


def set_batch_graph(on):
    """infx_set_batch_graph: batched launches (ScanBatch) through a captured
    HIP graph, replayed while the launch shape stays the same."""
    _lib.infx_set_batch_graph(1 if on else 0)


def last_call_timing():
    w, k, c, n = C.c_double(), C.c_double(), C.c_int64(), C.c_int()
    _lib.infx_last_call_timing(C.byref(w), C.byref(k), C.byref(c), C.byref(n))
    return {"wall_ms": w.value, "kernel_ms": k.value, "cells": c.value, "launches": n.value}


def digitize(seq):
    """Text sequence -> uint8 dsq[0..L+1] with sentinels (Easel convention)."""
    b = seq.encode() if isinstance(seq, str) else bytes(seq)
    dsq = np.zeros(len(b) + 2, dtype=np.uint8)
    st = _lib.infx_abc_Digitize(b, len(b), dsq.ctypes.data)
    if st != eslOK:
        raise ValueError("sequence has characters outside the RNA alphabet")
    return dsq


def dsq_from_codes(codes):
    """Residue codes (0..15) -> dsq with sentinels at 0 and L+1."""
    codes = np.asarray(codes, dtype=np.uint8)
    dsq = np.empty(len(codes) + 2, dtype=np.uint8)
    dsq[0] = dsq[-1] = 255
    dsq[1:-1] = codes
    return dsq


class CM:
    """A covariance model read from an INFERNAL1/a file (first model only)."""

    def __init__(self, ptr):
        self._p = ptr

    @classmethod
    def read(cls, path):
        p = _CMP()
        eb = C.create_string_buffer(ERRBUF)
        st = _lib.infx_cm_Read(os.fsencode(path), C.byref(p), eb)
        if st != eslOK:
            raise ScanError(st, eb.value.decode())
        return cls(p)

    @classmethod
    def from_text(cls, text):
        p = _CMP()
        eb = C.create_string_buffer(ERRBUF)
        st = _lib.infx_cm_ReadText(text.encode(), C.byref(p), eb)
        if st != eslOK:
            raise ScanError(st, eb.value.decode())
        return cls(p)

    def __del__(self):
        if getattr(self, "_p", None):
            _lib.FreeCM(self._p)
            self._p = None

    @property
    def ptr(self):
        return self._p

    @property
    def c(self):
        return self._p.contents

    def configure(self, local=True, trunc=False, scanmx=True, trscanmx=False, W=-1):
        c = self.c
        c.config_opts = ((CM_CONFIG_LOCAL if local else 0) | (CM_CONFIG_TRUNC if trunc else 0) |
                         (CM_CONFIG_SCANMX if scanmx else 0) | (CM_CONFIG_TRSCANMX if trscanmx else 0))
        eb = C.create_string_buffer(ERRBUF)
        st = _lib.cm_Configure(self._p, eb, int(W))
        if st != eslOK:
            raise ScanError(st, eb.value.decode())
        return self

    # convenience views
    M = property(lambda s: s.c.M)
    W = property(lambda s: s.c.W)
    clen = property(lambda s: s.c.clen)
    nodes = property(lambda s: s.c.nodes)

    def sttype(self):
        c = self.c
        return np.array([ord(c.sttype[v]) for v in range(c.M)], dtype=np.int32)

    def ints(self, name, n=None):
        arr = getattr(self.c, name)
        n = self.c.M if n is None else n
        return np.array([arr[i] for i in range(n)], dtype=np.int64)

    def floats(self, name, n=None):
        arr = getattr(self.c, name)
        n = self.c.M if n is None else n
        return np.array([arr[i] for i in range(n)], dtype=np.float32)

    def cells_per_res(self, qdbidx, truncated=False):
        cells, ops = C.c_double(), C.c_double()
        st = _lib.infx_CountSearchCells(self._p, qdbidx, 1 if truncated else 0, C.byref(cells), C.byref(ops))
        if st != eslOK:
            raise ScanError(st, "count failed")
        return cells.value, ops.value

    def work_per_res(self, qdbidx, truncated=False):
        """(cells, CYK ops, reductions) per residue at steady state: Inside ops
        count each reduction as one 8-op log-sum instead of one max."""
        cells, ops, red = C.c_double(), C.c_double(), C.c_double()
        st = _lib.infx_CountSearchWork(self._p, qdbidx, 1 if truncated else 0, C.byref(cells), C.byref(ops), C.byref(red))
        if st != eslOK:
            raise ScanError(st, "count failed")
        return cells.value, ops.value, red.value


def _hits_list(th):
    out = []
    t = th.contents
    for h in range(t.N):
        x = t.unsrt[h]
        out.append((x.start, x.stop, x.root, ord(x.mode) if isinstance(x.mode, bytes) else x.mode, x.score))
    return out


def _scan(fname, cm, qdbidx, dsq, i0, j0, cutoff, do_null3, env_cutoff, want_vsc, want_hits, pass_idx=None,
          inside=False, greedy=True):
    c = cm.c
    if inside:
        c.search_opts |= CM_SEARCH_INSIDE
    else:
        c.search_opts &= ~CM_SEARCH_INSIDE
    if greedy:
        c.search_opts &= ~CM_SEARCH_CMNOTGREEDY
    else:
        c.search_opts |= CM_SEARCH_CMNOTGREEDY
    dsq = np.ascontiguousarray(dsq, dtype=np.uint8)
    eb = C.create_string_buffer(ERRBUF)
    th = _lib.cm_tophits_Create() if want_hits else None
    envi, envj, sc = C.c_int64(), C.c_int64(), C.c_float()
    vsc = _fp()
    fn = getattr(_lib, fname)
    if pass_idx is None:
        st = fn(cm.ptr, eb, c.smx, qdbidx, dsq.ctypes.data, i0, j0, cutoff, th, int(do_null3), env_cutoff,
                C.byref(envi), C.byref(envj), C.byref(vsc) if want_vsc else None, C.byref(sc))
        mode = TRMODE_J
    else:
        md = C.create_string_buffer(1)
        st = fn(cm.ptr, eb, c.trsmx, qdbidx, pass_idx, dsq.ctypes.data, i0, j0, cutoff, th, int(do_null3), env_cutoff,
                C.byref(envi), C.byref(envj), C.byref(vsc) if want_vsc else None, md, C.byref(sc))
        mode = md.raw[0]
    if st != eslOK:
        if th:
            _lib.cm_tophits_Destroy(th)
        raise ScanError(st, eb.value.decode(errors="replace"))
    hits = _hits_list(th) if th else None
    if th:
        _lib.cm_tophits_Destroy(th)
    v = None
    if want_vsc:
        v = np.array([vsc[k] for k in range(c.M)], dtype=np.float32)
        C.CDLL(None).free(C.cast(vsc, C.c_void_p))
    return {"hits": hits, "envi": envi.value, "envj": envj.value, "sc": sc.value, "vsc": v, "mode": mode}


def ScanWindows(algo, cm, qdbidx, dsq, windows, cutoff, do_null3=False, greedy=True):
    """infx_FastScanWindows: many windows [(i, j), ...] (sorted, disjoint) of
    one target in one launch; algo "cyk", "finside" or "iinside".  The hits
    are those of one standard scan per window."""
    semi = {"cyk": 0, "finside": 1, "iinside": 2}[algo]
    c = cm.c
    if semi:
        c.search_opts |= CM_SEARCH_INSIDE
    else:
        c.search_opts &= ~CM_SEARCH_INSIDE
    if greedy:
        c.search_opts &= ~CM_SEARCH_CMNOTGREEDY
    else:
        c.search_opts |= CM_SEARCH_CMNOTGREEDY
    dsq = np.ascontiguousarray(dsq, dtype=np.uint8)
    wi = np.ascontiguousarray([w[0] for w in windows], dtype=np.int64)
    wj = np.ascontiguousarray([w[1] for w in windows], dtype=np.int64)
    eb = C.create_string_buffer(ERRBUF)
    th = _lib.cm_tophits_Create()
    sc = C.c_float()
    f = _lib.infx_FastScanWindows
    f.argtypes = [C.c_int, C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                  C.c_float, C.c_void_p, C.c_int, C.POINTER(C.c_float)]
    st = f(semi, C.cast(cm.ptr, C.c_void_p), eb, C.cast(c.smx, C.c_void_p), qdbidx, dsq.ctypes.data, wi.ctypes.data,
           wj.ctypes.data, len(windows), cutoff, th, int(do_null3), C.byref(sc))
    if st != eslOK:
        _lib.cm_tophits_Destroy(th)
        raise ScanError(st, eb.value.decode(errors="replace"))
    hits = _hits_list(th)
    _lib.cm_tophits_Destroy(th)
    return {"hits": hits, "sc": sc.value}


def ScanBatch(algo, cms, qdbidx, dsq, windows, cutoffs, greedy=True):
    """infx_FastScanBatch: several models' windows of one target in one
    launch; windows are (i, j, model index), each model's sorted and
    disjoint.  Returns per model {"hits": [...], "sc": best}, equal to
    ScanWindows over that model's windows."""
    semi = {"cyk": 0, "finside": 1, "iinside": 2}[algo]
    n = len(cms)
    if len(cutoffs) != n:
        raise ValueError("ScanBatch: one cutoff per model (%d models, %d cutoffs)" % (n, len(cutoffs)))
    for cm in cms:
        c = cm.c
        if semi:
            c.search_opts |= CM_SEARCH_INSIDE
        else:
            c.search_opts &= ~CM_SEARCH_INSIDE
        if greedy:
            c.search_opts &= ~CM_SEARCH_CMNOTGREEDY
        else:
            c.search_opts |= CM_SEARCH_CMNOTGREEDY
    dsq = np.ascontiguousarray(dsq, dtype=np.uint8)
    wi = np.ascontiguousarray([w[0] for w in windows], dtype=np.int64)
    wj = np.ascontiguousarray([w[1] for w in windows], dtype=np.int64)
    wm = np.ascontiguousarray([w[2] for w in windows], dtype=np.int32)
    cms_arr = (C.c_void_p * n)(*[C.cast(cm.ptr, C.c_void_p) for cm in cms])
    smx_arr = (C.c_void_p * n)(*[C.cast(cm.c.smx, C.c_void_p) for cm in cms])
    cut = np.ascontiguousarray(cutoffs, dtype=np.float32)
    ths = [_lib.cm_tophits_Create() for _ in range(n)]
    th_arr = (C.c_void_p * n)(*[C.cast(t, C.c_void_p) for t in ths])
    sc = np.zeros(n, np.float32)
    f = _lib.infx_FastScanBatch
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    eb = C.create_string_buffer(ERRBUF)
    st = f(semi, n, cms_arr, smx_arr, eb, qdbidx, dsq.ctypes.data, wi.ctypes.data, wj.ctypes.data, wm.ctypes.data,
           len(windows), cut.ctypes.data, th_arr, sc.ctypes.data)
    out = [{"hits": _hits_list(t), "sc": float(sc[m])} for m, t in enumerate(ths)] if st == eslOK else None
    for t in ths:
        _lib.cm_tophits_Destroy(t)
    if st != eslOK:
        raise ScanError(st, eb.value.decode(errors="replace"))
    return out


def TrScanWindows(algo, cm, qdbidx, pass_idx, dsq, windows, cutoff, do_null3=False, env_cutoff=0.0, want_hits=True,
                  greedy=True):
    """infx_TrScanWindows: one truncated scan per window [(i, j), ...] of one
    target, all in one launch (one workgroup per window); algo "cyk"
    (RefTrCYKScan) or "iinside" (RefITrInsideScan).  Per window: best score,
    its mode, envelope and (want_hits) hit list, as one call per window."""
    inside = {"cyk": 0, "iinside": 1}[algo]
    c = cm.c
    if inside:
        c.search_opts |= CM_SEARCH_INSIDE
    else:
        c.search_opts &= ~CM_SEARCH_INSIDE
    if greedy:
        c.search_opts &= ~CM_SEARCH_CMNOTGREEDY
    else:
        c.search_opts |= CM_SEARCH_CMNOTGREEDY
    dsq = np.ascontiguousarray(dsq, dtype=np.uint8)
    n = len(windows)
    wi = np.ascontiguousarray([w[0] for w in windows], dtype=np.int64)
    wj = np.ascontiguousarray([w[1] for w in windows], dtype=np.int64)
    ths = [_lib.cm_tophits_Create() for _ in range(n)] if want_hits else []
    arr = (C.c_void_p * n)(*[C.cast(t, C.c_void_p) for t in ths]) if want_hits else None
    sc = np.zeros(n, np.float32)
    md = np.zeros(n, np.int8)
    ei = np.zeros(n, np.int64)
    ej = np.zeros(n, np.int64)
    f = _lib.infx_TrScanWindows
    f.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_float, C.c_void_p, C.c_int, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    eb = C.create_string_buffer(ERRBUF)
    st = f(C.cast(cm.ptr, C.c_void_p), eb, C.cast(c.trsmx, C.c_void_p), inside, qdbidx, pass_idx, dsq.ctypes.data, n,
           wi.ctypes.data, wj.ctypes.data, cutoff, arr, int(do_null3), env_cutoff, sc.ctypes.data, md.ctypes.data,
           ei.ctypes.data, ej.ctypes.data)
    hits = [_hits_list(t) for t in ths] if (want_hits and st == eslOK) else None
    for t in ths:
        _lib.cm_tophits_Destroy(t)
    if st != eslOK:
        raise ScanError(st, eb.value.decode(errors="replace"))
    return {"sc": sc, "mode": md, "envi": ei, "envj": ej, "hits": hits}


def FastCYKScan(cm, qdbidx, dsq, i0, j0, cutoff, do_null3=False, env_cutoff=0.0, want_vsc=False, want_hits=True, greedy=True):
    """src/cm_dpsearch.c:90 FastCYKScan on the GPU."""
    return _scan("FastCYKScan", cm, qdbidx, dsq, i0, j0, cutoff, do_null3, env_cutoff, want_vsc, want_hits, greedy=greedy)


def FastFInsideScan(cm, qdbidx, dsq, i0, j0, cutoff, do_null3=False, env_cutoff=0.0, want_vsc=False, want_hits=True, greedy=True):
    """src/cm_dpsearch.c:1731 FastFInsideScan on the GPU."""
    return _scan("FastFInsideScan", cm, qdbidx, dsq, i0, j0, cutoff, do_null3, env_cutoff, want_vsc, want_hits,
                 inside=True, greedy=greedy)


def FastIInsideScan(cm, qdbidx, dsq, i0, j0, cutoff, do_null3=False, env_cutoff=0.0, want_vsc=False, want_hits=True, greedy=True):
    """src/cm_dpsearch.c:1106 FastIInsideScan on the GPU."""
    return _scan("FastIInsideScan", cm, qdbidx, dsq, i0, j0, cutoff, do_null3, env_cutoff, want_vsc, want_hits,
                 inside=True, greedy=greedy)


def RefTrCYKScan(cm, qdbidx, pass_idx, dsq, i0, j0, cutoff, do_null3=False, env_cutoff=0.0, want_vsc=False, want_hits=True, greedy=True):
    """src/cm_dpsearch_trunc.c:79 RefTrCYKScan on the GPU."""
    return _scan("RefTrCYKScan", cm, qdbidx, dsq, i0, j0, cutoff, do_null3, env_cutoff, want_vsc, want_hits,
                 pass_idx=pass_idx, greedy=greedy)


def RefITrInsideScan(cm, qdbidx, pass_idx, dsq, i0, j0, cutoff, do_null3=False, env_cutoff=0.0, want_vsc=False, want_hits=True, greedy=True):
    """src/cm_dpsearch_trunc.c:741 RefITrInsideScan on the GPU."""
    return _scan("RefITrInsideScan", cm, qdbidx, dsq, i0, j0, cutoff, do_null3, env_cutoff, want_vsc, want_hits,
                 pass_idx=pass_idx, inside=True, greedy=greedy)
