"""Learned-gate kernels (infx_gate_*) against plain PyTorch fp32 references
of the same operations on the same bf16-rounded inputs, computed on the CPU
(the kernels run on the library's own device buffers).

Tolerances: outputs stored in bf16 carry a relative rounding of 2^-8; fp32
accumulations over up to 1,152 products differ from torch's order by a few
ulps.  So bf16 outputs are checked to 1.6e-2 relative (of the layer's scale),
fp32 outputs to 1e-4 relative.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gate():
    from infernal_amd import gate
    return gate


def _bf(x):
    """float tensor -> bf16-rounded float tensor (CPU)"""
    return x.to(torch.bfloat16).float()


def _up(g, t, dt):
    a = t.detach().cpu().numpy()
    return g.DevBuf.upload(g.to_bf16_bits(a) if dt == "bf16" else a, dt)


def _down(b):
    from infernal_amd import gate as g
    a = b.get()
    return torch.from_numpy(g.from_bf16_bits(a) if b.dt == "bf16" else a.astype(np.float32))


def _ref_conv(x, w, cin, kp, seg, bias=None, relu=True, dil=1):
    """x [n][cin] float, w [128][kp] float (kk = k*cin + ci) -> [n][128]."""
    n = x.shape[0]
    wk = w[:, :9 * cin].reshape(128, 9, cin).permute(0, 2, 1)   # [co][ci][k]
    outs = []
    for a in range(0, n, seg):
        xs = x[a:a + seg].t().unsqueeze(0)                          # [1][cin][L]
        y = torch.nn.functional.conv1d(xs, wk, bias=bias, padding=4 * dil, dilation=dil)[0].t()
        outs.append(torch.relu(y) if relu else y)
    return torch.cat(outs)


def _close(got, ref, rel):
    scale = ref.abs().max().item() + 1e-6
    err = (got.float() - ref).abs().max().item()
    assert err <= rel * scale, (err, scale)


@pytest.mark.parametrize("cin,kp,n,seg,dil", [(4, 64, 1000, 1000, 1), (4, 64, 1024, 256, 3), (128, 1152, 777, 777, 1),
                                              (128, 1152, 1024, 128, 9), (128, 1152, 2000, 1000, 27)])
def test_conv_forward(cin, kp, n, seg, dil):
    g = _gate()
    print("case conv", cin, kp, n, seg, flush=True)
    gen = torch.Generator().manual_seed(cin + n)
    x = _bf(torch.randn(n, cin, generator=gen))
    w = torch.zeros(128, kp)
    w[:, :9 * cin] = torch.randn(128, 9 * cin, generator=gen) / math.sqrt(9 * cin)
    w = _bf(w)
    b = torch.randn(128, generator=gen)
    dx, dw, db = _up(g, x, "bf16"), _up(g, w, "bf16"), _up(g, b, "f32")
    y = g.DevBuf((n, 128), "bf16")
    g._call("infx_gate_conv", dx.ptr, n, seg, cin, dil, dw.ptr, kp, db.ptr, y.ptr, 0)
    _close(_down(y), _ref_conv(x, w, cin, kp, seg, bias=b, dil=dil), 1.6e-2)
    yf = g.DevBuf((n, 128), "f32")
    g._call("infx_gate_conv", dx.ptr, n, seg, cin, dil, dw.ptr, kp, None, yf.ptr, 2)
    _close(_down(yf), _ref_conv(x, w, cin, kp, seg, relu=False, dil=dil), 1e-4)


def test_onehot_and_gather():
    g = _gate()
    codes = np.array([0, 1, 2, 3, 4, 15, 2], dtype=np.uint8)
    dc = g.DevBuf.upload(codes, "u8")
    x = g.DevBuf((7, 4), "bf16")
    g._call("infx_gate_onehot", dc.ptr, 7, x.ptr)
    ref = np.zeros((7, 4), dtype=np.float32)
    for i, c in enumerate(codes):
        if c < 4:
            ref[i, c] = 1
    assert np.array_equal(g.from_bf16_bits(x.get()), ref)
    big = np.arange(1000, dtype=np.int64) % 251
    lab = (np.arange(1000) % 7).astype(np.float32)
    st = np.array([5, 900, 0], dtype=np.int64)
    oc, ol = g.DevBuf(3 * 64, "u8"), g.DevBuf(3 * 64, "f32")
    # every buffer a kernel reads stays referenced until the call returns
    db, dl, ds = g.DevBuf.upload(big, "u8"), g.DevBuf.upload(lab, "f32"), g.DevBuf.upload(st, "i64")
    g._call("infx_gate_gather", db.ptr, dl.ptr, ds.ptr, 3, 64, oc.ptr, ol.ptr)
    want = np.concatenate([np.arange(s, s + 64) for s in st])
    assert np.array_equal(oc.get(), big[want].astype(np.uint8)) and np.array_equal(ol.get(), lab[want])


@pytest.mark.parametrize("cin,kp,n,seg,dil", [(4, 64, 3000, 1000, 1), (128, 1152, 2048, 512, 1), (128, 1152, 2048, 1024, 9)])
def test_weight_and_data_gradients(cin, kp, n, seg, dil):
    g = _gate()
    print("case grads", cin, kp, n, seg, flush=True)
    gen = torch.Generator().manual_seed(7 + cin)
    x = _bf(torch.randn(n, cin, generator=gen))
    dy = _bf(torch.randn(n, 128, generator=gen))
    w = torch.zeros(128, kp)
    w[:, :9 * cin] = torch.randn(128, 9 * cin, generator=gen) / math.sqrt(9 * cin)
    xr = x.clone().requires_grad_(True)
    wr = _bf(w).clone().requires_grad_(True)
    (_ref_conv(xr, wr, cin, kp, seg, relu=False, dil=dil) * dy).sum().backward()
    ddy, dxb = _up(g, dy, "bf16"), _up(g, x, "bf16")
    dw = g.DevBuf.zeros((128, kp), "f32")
    g._call("infx_gate_wgrad", ddy.ptr, dxb.ptr, n, seg, cin, dil, kp, dw.ptr)
    got = _down(dw)
    _close(got, wr.grad, 1e-4)
    if 9 * cin < kp:
        assert got[:, 9 * cin:].abs().max().item() == 0.0
    if cin == 128:
        wm = _up(g, w, "f32")
        wf = g.DevBuf.zeros((128, kp), "bf16")
        wd = g.DevBuf.zeros((128, kp), "bf16")
        g._call("infx_gate_pack", wm.ptr, kp, cin, wf.ptr, wd.ptr)
        assert torch.equal(_down(wf), _bf(w))
        dxo = g.DevBuf((n, 128), "bf16")
        g._call("infx_gate_conv", ddy.ptr, n, seg, 128, dil, wd.ptr, kp, None, dxo.ptr, 1)
        _close(_down(dxo), xr.grad, 1.6e-2)
    cs = g.DevBuf.zeros(128, "f32")
    g._call("infx_gate_colsum", ddy.ptr, n, cs.ptr)
    _close(_down(cs), dy.sum(0), 1e-4)


def test_head_loss_relu_adam():
    g = _gate()
    n = 5000
    gen = torch.Generator().manual_seed(3)
    y = _bf(torch.relu(torch.randn(n, 128, generator=gen)))
    hw = torch.randn(128, generator=gen)
    hb = torch.randn(1, generator=gen)
    dyb, dhw_in, dhb_in = _up(g, y, "bf16"), _up(g, hw, "f32"), _up(g, hb, "f32")
    s = g.DevBuf(n, "f32")
    g._call("infx_gate_head", dyb.ptr, n, dhw_in.ptr, dhb_in.ptr, s.ptr)
    st = _down(s)
    _close(st, y @ hw + hb, 1e-4)
    lab = (torch.rand(n, generator=gen) < 0.2).float()
    pw = 3.0
    dy = g.DevBuf((n, 128), "bf16")
    dhw, dhb, loss = g.DevBuf.zeros(128, "f32"), g.DevBuf.zeros(1, "f32"), g.DevBuf.zeros(1, "f32")
    dlab = _up(g, lab, "f32")
    g._call("infx_gate_head_bwd", s.ptr, dlab.ptr, dyb.ptr, n, dhw_in.ptr, pw, dy.ptr,
            dhw.ptr, dhb.ptr, loss.ptr)
    sr = st.clone().requires_grad_(True)
    wgt = torch.where(lab > 0.5, torch.full_like(lab, pw), torch.ones_like(lab))
    lref = (wgt * torch.nn.functional.binary_cross_entropy_with_logits(sr, lab, reduction="none")).mean()
    lref.backward()
    assert abs(_down(loss).item() - lref.item()) <= 1e-4 * max(1.0, lref.item())
    _close(_down(dhw), y.t() @ sr.grad, 1e-3)
    assert abs(_down(dhb).item() - sr.grad.sum().item()) <= 1e-4
    dyt = _down(dy)
    _close(dyt, sr.grad[:, None] * hw[None, :], 1.6e-2)
    # ReLU backward: keeps the gradient where the activation is positive
    g._call("infx_gate_relu_bwd", dy.ptr, dyb.ptr, n * 128)
    assert torch.equal(_down(dy), torch.where(y > 0, dyt, torch.zeros_like(dyt)))
    # Adam, first step
    w0 = torch.randn(1000, generator=gen)
    gr = torch.randn(1000, generator=gen)
    w, m, v = _up(g, w0, "f32"), g.DevBuf.zeros(1000, "f32"), g.DevBuf.zeros(1000, "f32")
    dgr = _up(g, gr, "f32")
    g._call("infx_gate_adam", w.ptr, dgr.ptr, m.ptr, v.ptr, 1000, 0.01, 0.9, 0.999, 1e-8, 1)
    mref, vref = 0.1 * gr, 0.001 * gr * gr
    wref = w0 - 0.01 * (mref / 0.1) / (torch.sqrt(vref / 0.001) + 1e-8)
    _close(_down(w), wref, 1e-5)


def test_training_learns_a_motif():
    """A planted 12-nt motif in random sequence: after 150 steps the gate
    ranks motif residues above the rest (AUC > 0.95)."""
    g = _gate()
    rng = np.random.default_rng(0)
    motif = rng.integers(0, 4, 12).astype(np.uint8)
    n = 200_000
    codes = rng.integers(0, 4, n).astype(np.uint8)
    lab = np.zeros(n, dtype=np.float32)
    for p in rng.integers(0, n - 12, 600):
        codes[p:p + 12] = motif
        lab[p:p + 12] = 1.0
    net = g.GateNet(seed=1)
    losses = g.train(net, codes[:150_000], lab[:150_000], 150, seg=512, nseg=16, lr=2e-3, pos_weight=4.0)
    s = net.score(codes[150_000:])
    t = lab[150_000:]
    pos, neg = s[t > 0.5], s[t < 0.5]
    auc = (pos[:, None] > neg[None, ::50]).mean()
    print("loss", losses[0], "->", losses[-1], "auc", auc)
    assert losses[-1] < 0.5 * losses[0]
    assert auc > 0.95
