"""The learned window gate (no counterpart in the reference).

A per-model 1-D convolutional scorer over one-hot ACGU decides which
stretches of a target go to the exact Inside scan:

  GateNet        4 convolutions (width 9, 128 channels, ReLU) and a
                 per-position linear head giving a logit per residue; bf16
                 weights and activations on the matrix cores, fp32
                 accumulation; trained with Adam on a logistic loss by the
                 library's own kernels (infx_gate_*)
  labels_from_hits   residues inside the plain method's hits are positives
  select_windows     residues whose logit passes the threshold, widened by
                     W-1 on each side (a hit of length <= W around them fits)
                     and merged, as the pipeline widens CYK survivors
  gated_inside       exact FastFInsideScan on those windows only

The plain scan (CYK and Inside over everything) stays the fallback and the
judge: recall is the fraction of the plain method's hits that the gated scan
returns.  PyTorch only holds device memory here; every computation on it is
a kernel of the library.
"""
import ctypes as C
import math

import numpy as np

from . import _lib, FastFInsideScan, dsq_from_codes

KP = (64, 1152, 1152, 1152)
CIN = (4, 128, 128, 128)
HALO = 16          # receptive-field radius of the 4 convolutions (4 x 4)

_vp, _i64, _int, _f = C.c_void_p, C.c_int64, C.c_int, C.c_float
_SIGS = {
    "infx_gate_onehot": [_vp, _i64, _vp],
    "infx_gate_conv": [_vp, _i64, _int, _int, _vp, _int, _vp, _vp, _int],
    "infx_gate_wgrad": [_vp, _vp, _i64, _int, _int, _int, _vp],
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


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class GateNet:
    """Weights (fp32 masters, bf16 copies) and Adam state on the GPU."""

    def __init__(self, seed=0):
        import torch
        self.torch = torch
        g = torch.Generator().manual_seed(seed)
        dev = "cuda"
        self.W, self.b = [], []
        for l in range(4):
            w = torch.zeros(128, KP[l])
            fan = 9 * CIN[l]
            w[:, :fan] = torch.randn(128, fan, generator=g) * math.sqrt(2.0 / fan)
            self.W.append(w.to(dev))
            self.b.append(torch.zeros(128, device=dev))
        self.hw = (torch.randn(128, generator=g) * math.sqrt(1.0 / 128)).to(dev)
        self.hb = torch.zeros(1, device=dev)
        self.Wf = [torch.zeros(128, KP[l], dtype=torch.bfloat16, device=dev) for l in range(4)]
        self.Wd = [None] + [torch.zeros(128, KP[l], dtype=torch.bfloat16, device=dev) for l in range(1, 4)]
        self.params = self.W + self.b + [self.hw, self.hb]
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.step = 0
        self.pack()

    def pack(self):
        for l in range(4):
            _call("infx_gate_pack", _p(self.W[l]), KP[l], CIN[l], _p(self.Wf[l]), _p(self.Wd[l]))

    def forward(self, codes, seg):
        """codes: uint8 device tensor (n residues, a multiple of seg).
        Returns (logits fp32 [n], activations [x, y1..y4])."""
        torch = self.torch
        n = codes.numel()
        x = torch.empty(n, 4, dtype=torch.bfloat16, device="cuda")
        _call("infx_gate_onehot", _p(codes), n, _p(x))
        acts = [x]
        for l in range(4):
            y = torch.empty(n, 128, dtype=torch.bfloat16, device="cuda")
            _call("infx_gate_conv", _p(acts[-1]), n, seg, CIN[l], _p(self.Wf[l]), KP[l], _p(self.b[l]), _p(y), 0)
            acts.append(y)
        s = torch.empty(n, dtype=torch.float32, device="cuda")
        _call("infx_gate_head", _p(acts[-1]), n, _p(self.hw), _p(self.hb), _p(s))
        return s, acts

    def train_step(self, codes, labels, seg, lr=1e-3, pos_weight=1.0):
        """One Adam step on a batch of segments; returns the mean loss."""
        torch = self.torch
        n = codes.numel()
        s, acts = self.forward(codes, seg)
        grads = [torch.zeros_like(p) for p in self.params]
        gW, gb, ghw, ghb = grads[0:4], grads[4:8], grads[8], grads[9]
        loss = torch.zeros(1, device="cuda")
        dy = torch.empty(n, 128, dtype=torch.bfloat16, device="cuda")
        _call("infx_gate_head_bwd", _p(s), _p(labels), _p(acts[4]), n, _p(self.hw), float(pos_weight),
              _p(dy), _p(ghw), _p(ghb), _p(loss))
        for l in range(3, -1, -1):
            _call("infx_gate_relu_bwd", _p(dy), _p(acts[l + 1]), n * 128)
            _call("infx_gate_wgrad", _p(dy), _p(acts[l]), n, seg, CIN[l], KP[l], _p(gW[l]))
            _call("infx_gate_colsum", _p(dy), n, _p(gb[l]))
            if l > 0:
                dx = torch.empty(n, 128, dtype=torch.bfloat16, device="cuda")
                _call("infx_gate_conv", _p(dy), n, seg, 128, _p(self.Wd[l]), KP[l], None, _p(dx), 1)
                dy = dx
        self.step += 1
        for p, g, m, v in zip(self.params, grads, self.m, self.v):
            _call("infx_gate_adam", _p(p), _p(g), _p(m), _p(v), p.numel(), float(lr), 0.9, 0.999, 1e-8, self.step)
        self.pack()
        return float(loss.item())

    def score(self, codes_np, chunk=1 << 22):
        """Logit per residue of a whole target (numpy uint8 codes), in
        overlapping device chunks (halo of the receptive field)."""
        torch = self.torch
        n = len(codes_np)
        out = np.empty(n, dtype=np.float32)
        dev = torch.from_numpy(np.ascontiguousarray(codes_np)).to("cuda")
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            lo, hi = max(0, a - HALO), min(n, b + HALO)
            s, _ = self.forward(dev[lo:hi].contiguous(), hi - lo)
            out[a:b] = s[a - lo:b - lo].cpu().numpy()
        return out


def labels_from_hits(n, hits):
    """1.0 at residues inside a hit (start..stop, 1-based inclusive)."""
    lab = np.zeros(n, dtype=np.float32)
    for h in hits:
        lo, hi = min(h[0], h[1]), max(h[0], h[1])
        lab[lo - 1:hi] = 1.0
    return lab


def select_windows(logits, tau, W):
    """Windows (1-based, inclusive) around residues with logit > tau."""
    n = len(logits)
    idx = np.nonzero(logits > tau)[0] + 1
    if len(idx) == 0:
        return []
    lo = np.maximum(1, idx - (W - 1))
    hi = np.minimum(n, idx + (W - 1))
    wins = []
    cur_lo, cur_hi = int(lo[0]), int(hi[0])
    for a, b in zip(lo[1:], hi[1:]):
        if a <= cur_hi + 1:
            cur_hi = max(cur_hi, int(b))
        else:
            wins.append((cur_lo, cur_hi))
            cur_lo, cur_hi = int(a), int(b)
    wins.append((cur_lo, cur_hi))
    return wins


def gated_inside(cm, dsq, windows, cutoff, qdbidx=2):
    """Exact FastFInsideScan over each window; hits (start, stop, score)."""
    hits = []
    for (i, j) in windows:
        r = FastFInsideScan(cm, qdbidx, dsq, i, j, cutoff, want_hits=True)
        hits.extend((h[0], h[1], h[4]) for h in r["hits"])
    return hits


def recall(plain_hits, gated_hits):
    """Fraction of plain hits that a gated hit overlaps by at least half."""
    if not plain_hits:
        return 1.0
    g = sorted((min(a, b), max(a, b)) for a, b, _ in gated_hits)
    starts = np.array([x[0] for x in g]) if g else np.zeros(0)
    found = 0
    for h in plain_hits:
        lo, hi = min(h[0], h[1]), max(h[0], h[1])
        k = np.searchsorted(starts, hi, side="right")
        ok = False
        for m in range(max(0, k - 8), k):
            glo, ghi = g[m]
            if min(hi, ghi) - max(lo, glo) + 1 >= 0.5 * (hi - lo + 1):
                ok = True
                break
        found += ok
    return found / len(plain_hits)
