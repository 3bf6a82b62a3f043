"""The learned window gate (no counterpart in the reference).

A per-model 1-D convolutional scorer over one-hot ACGU decides which
stretches of a target go to the exact Inside scan:

  GateNet        4 convolutions (width 9, dilations 1/3/9/27, 128 channels,
                 ReLU) and a
                 per-position linear head giving a logit per residue; bf16
                 weights and activations on the matrix cores, fp32
                 accumulation; trained with Adam on a logistic loss by the
                 library's own kernels (infx_gate_*)
  labels_from_hits   residues inside the plain method's hits are positives
  smooth             the logits averaged over 64 residues
  select_windows     residues whose smoothed logit passes the threshold, widened by
                     W-1 on each side (a hit of length <= W around them fits)
                     and merged, as the pipeline widens CYK survivors, or by
                     the widening fit_widen sets on held-out hits
  gated_inside       exact FastFInsideScan on those windows only

The plain scan (CYK and Inside over everything) stays the fallback and the
judge: recall is the fraction of the plain method's hits that the gated scan
returns.  Device memory comes from the library's own runtime (DevBuf), so
the gate's kernels, buffers and copies are ordered on one null stream.
"""
import ctypes as C
import math

import numpy as np

from . import _lib, ScanWindows

KP = (64, 1152, 1152, 1152)
CIN = (4, 128, 128, 128)
DIL = (1, 3, 9, 27)              # dilations: a 321-residue receptive field
HALO = 4 * sum(DIL)              # its radius

_vp, _i64, _int, _f, _sz = C.c_void_p, C.c_int64, C.c_int, C.c_float, C.c_size_t
_SIGS = {
    "infx_dev_malloc": [_sz, C.POINTER(C.c_void_p)],
    "infx_dev_free": [_vp],
    "infx_dev_memset": [_vp, _int, _sz],
    "infx_dev_upload": [_vp, _vp, _sz],
    "infx_dev_download": [_vp, _vp, _sz],
    "infx_dev_sync": [],
    "infx_gate_gather": [_vp, _vp, _vp, _int, _int, _vp, _vp],
    "infx_gate_onehot": [_vp, _i64, _vp],
    "infx_gate_conv": [_vp, _i64, _int, _int, _int, _vp, _int, _vp, _vp, _int],
    "infx_gate_wgrad": [_vp, _vp, _i64, _int, _int, _int, _int, _vp],
    "infx_gate_relu_bwd": [_vp, _vp, _i64],
    "infx_gate_colsum": [_vp, _i64, _vp],
    "infx_gate_head": [_vp, _i64, _vp, _vp, _vp],
    "infx_gate_head_bwd": [_vp, _vp, _vp, _i64, _vp, _f, _vp, _vp, _vp, _vp],
    "infx_gate_adam": [_vp, _vp, _vp, _vp, _i64, _f, _f, _f, _f, _int],
    "infx_gate_pack": [_vp, _int, _int, _vp, _vp],
}
for _name, _args in _SIGS.items():
    getattr(_lib, _name).argtypes = _args
    getattr(_lib, _name).restype = _int


def _call(name, *args):
    st = getattr(_lib, name)(*args)
    if st != 0:
        raise RuntimeError("%s failed: HIP error %d" % (name, st))


# numpy dtypes of the device arrays; bf16 travels as its 16-bit pattern
_DT = {"u8": np.uint8, "f32": np.float32, "bf16": np.uint16, "i64": np.int64}


def to_bf16_bits(a):
    """float32 array -> bf16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class DevBuf:
    """A device array (shape, dtype tag) from the library's runtime."""

    def __init__(self, shape, dt):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dt = dt
        self.n = int(np.prod(self.shape))
        self.nbytes = self.n * np.dtype(_DT[dt]).itemsize
        p = C.c_void_p()
        _call("infx_dev_malloc", self.nbytes, C.byref(p))
        self.ptr = p

    @classmethod
    def zeros(cls, shape, dt):
        b = cls(shape, dt)
        _call("infx_dev_memset", b.ptr, 0, b.nbytes)
        return b

    @classmethod
    def upload(cls, arr, dt):
        a = np.ascontiguousarray(arr, dtype=_DT[dt])
        b = cls(a.shape, dt)
        _call("infx_dev_upload", b.ptr, a.ctypes.data_as(C.c_void_p), b.nbytes)
        return b

    def put(self, arr):
        a = np.ascontiguousarray(arr, dtype=_DT[self.dt])
        assert a.size == self.n
        _call("infx_dev_upload", self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes)

    def get(self):
        a = np.empty(self.shape, dtype=_DT[self.dt])
        _call("infx_dev_download", a.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes)
        return a

    def zero(self):
        _call("infx_dev_memset", self.ptr, 0, self.nbytes)

    def __del__(self):
        try:
            if getattr(self, "ptr", None) is not None and self.ptr.value:
                _lib.infx_dev_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


class GateNet:
    """Weights (fp32 masters, bf16 copies) and Adam state on the GPU."""

    def __init__(self, seed=0):
        rng = np.random.default_rng(seed)
        self.W, self.b = [], []
        for l in range(4):
            w = np.zeros((128, KP[l]), dtype=np.float32)
            fan = 9 * CIN[l]
            w[:, :fan] = rng.standard_normal((128, fan)) * math.sqrt(2.0 / fan)
            self.W.append(DevBuf.upload(w, "f32"))
            self.b.append(DevBuf.zeros(128, "f32"))
        self.hw = DevBuf.upload(rng.standard_normal(128) * math.sqrt(1.0 / 128), "f32")
        self.hb = DevBuf.zeros(1, "f32")
        self.Wf = [DevBuf.zeros((128, KP[l]), "bf16") for l in range(4)]
        self.Wd = [None] + [DevBuf.zeros((128, KP[l]), "bf16") for l in range(1, 4)]
        self.params = self.W + self.b + [self.hw, self.hb]
        self.grads = [DevBuf.zeros(p.shape, "f32") for p in self.params]
        self.m = [DevBuf.zeros(p.shape, "f32") for p in self.params]
        self.v = [DevBuf.zeros(p.shape, "f32") for p in self.params]
        self.step = 0
        self._bufs = {}
        self.pack()

    def pack(self):
        for l in range(4):
            _call("infx_gate_pack", self.W[l].ptr, KP[l], CIN[l], self.Wf[l].ptr, self.Wd[l].ptr if self.Wd[l] else None)

    def _buf(self, key, shape, dt):
        b = self._bufs.get(key)
        if b is None or b.shape != tuple(shape):
            b = DevBuf(shape, dt)
            self._bufs[key] = b
        return b

    def forward(self, codes, n, seg):
        """codes: DevBuf (u8) of n residues (a multiple of seg).  Returns
        (logits DevBuf [n] fp32, activations [x, y1..y4])."""
        x = self._buf("x", (n, 4), "bf16")
        _call("infx_gate_onehot", codes.ptr, n, x.ptr)
        acts = [x]
        for l in range(4):
            y = self._buf("y%d" % l, (n, 128), "bf16")
            _call("infx_gate_conv", acts[-1].ptr, n, seg, CIN[l], DIL[l], self.Wf[l].ptr, KP[l], self.b[l].ptr, y.ptr, 0)
            acts.append(y)
        s = self._buf("s", (n,), "f32")
        _call("infx_gate_head", acts[-1].ptr, n, self.hw.ptr, self.hb.ptr, s.ptr)
        return s, acts

    def train_step(self, codes, labels, n, seg, lr=1e-3, pos_weight=1.0):
        """One Adam step on a batch (device codes and labels of n residues,
        segments of seg); returns the mean loss."""
        s, acts = self.forward(codes, n, seg)
        for g in self.grads:
            g.zero()
        gW, gb, ghw, ghb = self.grads[0:4], self.grads[4:8], self.grads[8], self.grads[9]
        loss = self._buf("loss", (1,), "f32")
        loss.zero()
        dy = self._buf("dy0", (n, 128), "bf16")
        dx = self._buf("dy1", (n, 128), "bf16")
        _call("infx_gate_head_bwd", s.ptr, labels.ptr, acts[4].ptr, n, self.hw.ptr, float(pos_weight),
              dy.ptr, ghw.ptr, ghb.ptr, loss.ptr)
        for l in range(3, -1, -1):
            _call("infx_gate_relu_bwd", dy.ptr, acts[l + 1].ptr, n * 128)
            _call("infx_gate_wgrad", dy.ptr, acts[l].ptr, n, seg, CIN[l], DIL[l], KP[l], gW[l].ptr)
            _call("infx_gate_colsum", dy.ptr, n, gb[l].ptr)
            if l > 0:
                _call("infx_gate_conv", dy.ptr, n, seg, 128, DIL[l], self.Wd[l].ptr, KP[l], None, dx.ptr, 1)
                dy, dx = dx, dy
        self.step += 1
        for p, g, m, v in zip(self.params, self.grads, self.m, self.v):
            _call("infx_gate_adam", p.ptr, g.ptr, m.ptr, v.ptr, p.n, float(lr), 0.9, 0.999, 1e-8, self.step)
        self.pack()
        return float(loss.get()[0])

    def score(self, codes_np, chunk=1 << 22):
        """Logit per residue of a whole target (numpy uint8 codes), in
        overlapping device chunks (halo of the receptive field)."""
        n = len(codes_np)
        out = np.empty(n, dtype=np.float32)
        codes_np = np.ascontiguousarray(codes_np, dtype=np.uint8)
        dc = self._buf("score_codes", (min(n, chunk + 2 * HALO),), "u8")
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            lo, hi = max(0, a - HALO), min(n, b + HALO)
            m = hi - lo
            if dc.n != m:
                dc = self._buf("score_codes", (m,), "u8")
            dc.put(codes_np[lo:hi])
            s, _ = self.forward(dc, m, m)
            out[a:b] = s.get()[a - lo:b - lo]
        return out


def train(net, codes_np, labels_np, steps, seg=1024, nseg=32, lr=2e-3, pos_weight=1.0, seed=5):
    """Adam steps on random segments of a device-resident target."""
    rng = np.random.default_rng(seed)
    n = len(codes_np)
    dcodes = DevBuf.upload(codes_np, "u8")
    dlab = DevBuf.upload(labels_np, "f32")
    starts = DevBuf(nseg, "i64")
    bc = DevBuf(nseg * seg, "u8")
    bl = DevBuf(nseg * seg, "f32")
    losses = []
    for _ in range(steps):
        starts.put(rng.integers(0, n - seg, nseg).astype(np.int64))
        _call("infx_gate_gather", dcodes.ptr, dlab.ptr, starts.ptr, nseg, seg, bc.ptr, bl.ptr)
        losses.append(net.train_step(bc, bl, nseg * seg, seg, lr=lr, pos_weight=pos_weight))
    return losses


def smooth(logits, k=64):
    """Mean of the logits over the k residues centred on each position (a
    homolog gives a run of high logits, background a scatter of single
    ones); the window is cut at the target ends."""
    x = np.asarray(logits, dtype=np.float64)
    n = len(x)
    c = np.concatenate([[0.0], np.cumsum(x)])
    lo = np.clip(np.arange(n) - k // 2, 0, n)
    hi = np.clip(np.arange(n) + (k - k // 2), 0, n)
    return ((c[hi] - c[lo]) / np.maximum(hi - lo, 1)).astype(np.float32)


def labels_from_hits(n, hits):
    """1.0 at residues inside a hit (start..stop, 1-based inclusive)."""
    lab = np.zeros(n, dtype=np.float32)
    for h in hits:
        lo, hi = min(h[0], h[1]), max(h[0], h[1])
        lab[lo - 1:hi] = 1.0
    return lab


def select_windows(logits, tau, W, widen=None):
    """Windows (1-based, inclusive) around residues with logit > tau, each
    widened by `widen` residues on each side (default W - 1: any hit of
    length <= W that touches a flagged residue fits) and merged."""
    n = len(logits)
    m = W - 1 if widen is None else int(min(max(widen, 0), W - 1))
    idx = np.nonzero(logits > tau)[0] + 1
    if len(idx) == 0:
        return []
    lo = np.maximum(1, idx - m)
    hi = np.minimum(n, idx + m)
    # a new window starts where a residue's window does not touch the previous one
    brk = np.nonzero(lo[1:] > np.maximum.accumulate(hi)[:-1] + 1)[0] + 1
    starts = np.concatenate([[0], brk])
    ends = np.concatenate([brk, [len(idx)]])
    hmax = np.maximum.accumulate(hi)
    return [(int(lo[a]), int(hmax[b - 1])) for a, b in zip(starts, ends)]


def windows_cover(windows, hits):
    """For each hit (start, stop, ...), whether one window holds all of it."""
    if not windows:
        return np.zeros(len(hits), dtype=bool)
    ws = np.array([a for a, _ in windows], dtype=np.int64)
    we = np.array([b for _, b in windows], dtype=np.int64)
    out = np.zeros(len(hits), dtype=bool)
    for i, h in enumerate(hits):
        lo, hi = min(h[0], h[1]), max(h[0], h[1])
        k = int(np.searchsorted(ws, lo, side="right")) - 1
        out[i] = k >= 0 and we[k] >= hi
    return out


def fit_widen(logits, tau, hits, W, slack=0.002):
    """The widening the gate's windows need, set on held-out hits as the
    threshold is: the smallest widening in 0 .. W - 1 whose windows hold
    whole as many held-out hits as W - 1 does, less `slack` of them (found
    by bisection: a wider window holds every hit a narrower one holds).
    The gate's labels are the residues inside hits, so the flagged residues
    already span most of a hit; the widening covers the ends the smoothed
    logits miss."""
    if not hits:
        return W - 1
    full = windows_cover(select_windows(logits, tau, W), hits).mean()
    lo, hi = 0, W - 1          # the answer lies in [lo, hi]; hi always qualifies
    while lo < hi:
        m = (lo + hi) // 2
        if windows_cover(select_windows(logits, tau, W, m), hits).mean() >= full - slack:
            hi = m
        else:
            lo = m + 1
    return hi


def gated_inside(cm, dsq, windows, cutoff, qdbidx=2):
    """Exact FastFInsideScan semantics on each window, all windows in one
    launch; hits (start, stop, score)."""
    if not windows:
        return []
    r = ScanWindows("finside", cm, qdbidx, dsq, windows, cutoff)
    return [(h[0], h[1], h[4]) for h in r["hits"]]


def recall(plain_hits, gated_hits):
    """Fraction of plain hits that a gated hit overlaps by at least half."""
    if not plain_hits:
        return 1.0
    g = sorted((min(a, b), max(a, b)) for a, b, _ in gated_hits)
    gs = np.array([x[0] for x in g], dtype=np.int64)
    ge = np.maximum.accumulate(np.array([x[1] for x in g], dtype=np.int64)) if g else gs
    found = 0
    for h in plain_hits:
        lo, hi = min(h[0], h[1]), max(h[0], h[1])
        k = int(np.searchsorted(gs, hi, side="right"))
        ok = False
        m = k - 1
        while m >= 0 and ge[m] >= lo:
            glo, ghi = g[m]
            if min(hi, ghi) - max(lo, glo) + 1 >= 0.5 * (hi - lo + 1):
                ok = True
                break
            m -= 1
        found += ok
    return found / len(plain_hits)
