"""Learned-gate kernels (infx_gate_*) against plain PyTorch fp32 references
of the same operations on the same bf16-rounded inputs.

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
    return x.to(torch.bfloat16)


def _ref_conv(x, w, cin, kp, seg, bias=None, relu=True):
    """x [n][cin] float, w [128][kp] float (kk = k*cin + ci) -> [n][128]."""
    n = x.shape[0]
    wk = w[:, :9 * cin].reshape(128, 9, cin).permute(0, 2, 1)   # [co][ci][k]
    out = torch.empty(n, 128, device=x.device)
    for a in range(0, n, seg):
        xs = x[a:a + seg].t().unsqueeze(0)                          # [1][cin][L]
        y = torch.nn.functional.conv1d(xs, wk, bias=bias, padding=4)[0].t()
        out[a:a + seg] = torch.relu(y) if relu else y
    return out


def _close(got, ref, rel):
    scale = ref.abs().max().item() + 1e-6
    err = (got.float() - ref).abs().max().item()
    assert err <= rel * scale, (err, scale)


@pytest.mark.parametrize("cin,kp,n,seg", [(4, 64, 1000, 1000), (4, 64, 1024, 256), (128, 1152, 777, 777), (128, 1152, 1024, 128)])
def test_conv_forward(cin, kp, n, seg):
    gate = _gate()
    print("case conv", cin, kp, n, seg, flush=True)
    g = torch.Generator().manual_seed(cin + n)
    x = _bf(torch.randn(n, cin, generator=g)).cuda()
    w = torch.zeros(128, kp)
    w[:, :9 * cin] = torch.randn(128, 9 * cin, generator=g) / math.sqrt(9 * cin)
    wb = _bf(w).cuda()
    b = torch.randn(128, generator=g).cuda()
    y = torch.empty(n, 128, dtype=torch.bfloat16, device="cuda")
    gate._call("infx_gate_conv", gate._p(x), n, seg, cin, gate._p(wb), kp, gate._p(b), gate._p(y), 0)
    ref = _ref_conv(x.float(), wb.float(), cin, kp, seg, bias=b)
    _close(y, ref, 1.6e-2)
    yf = torch.empty(n, 128, dtype=torch.float32, device="cuda")
    gate._call("infx_gate_conv", gate._p(x), n, seg, cin, gate._p(wb), kp, None, gate._p(yf), 2)
    _close(yf, _ref_conv(x.float(), wb.float(), cin, kp, seg, relu=False), 1e-4)


def test_onehot():
    gate = _gate()
    codes = torch.tensor([0, 1, 2, 3, 4, 15, 2], dtype=torch.uint8, device="cuda")
    x = torch.empty(7, 4, dtype=torch.bfloat16, device="cuda")
    gate._call("infx_gate_onehot", gate._p(codes), 7, gate._p(x))
    ref = torch.zeros(7, 4)
    for i, c in enumerate([0, 1, 2, 3, 4, 15, 2]):
        if c < 4:
            ref[i, c] = 1
    assert torch.equal(x.float().cpu(), ref)


@pytest.mark.parametrize("cin,kp,n,seg", [(4, 64, 3000, 1000), (128, 1152, 2048, 512)])
def test_weight_and_data_gradients(cin, kp, n, seg):
    gate = _gate()
    print("case grads", cin, kp, n, seg, flush=True)
    g = torch.Generator().manual_seed(7 + cin)
    x = _bf(torch.randn(n, cin, generator=g)).cuda()
    dy = _bf(torch.randn(n, 128, generator=g)).cuda()
    w = torch.zeros(128, kp)
    w[:, :9 * cin] = torch.randn(128, 9 * cin, generator=g) / math.sqrt(9 * cin)
    w = w.cuda()
    # autograd reference on the same bf16 values
    xr = x.float().clone().requires_grad_(True)
    wr = _bf(w).float().clone().requires_grad_(True)
    yr = _ref_conv(xr, wr, cin, kp, seg, relu=False)
    (yr * dy.float()).sum().backward()
    dw = torch.zeros(128, kp, device="cuda")
    gate._call("infx_gate_wgrad", gate._p(dy), gate._p(x), n, seg, cin, kp, gate._p(dw))
    _close(dw, wr.grad, 1e-4)
    if 9 * cin < kp:
        assert dw[:, 9 * cin:].abs().max().item() == 0.0
    if cin == 128:
        wf = torch.zeros(128, kp, dtype=torch.bfloat16, device="cuda")
        wd = torch.zeros(128, kp, dtype=torch.bfloat16, device="cuda")
        gate._call("infx_gate_pack", gate._p(w), kp, cin, gate._p(wf), gate._p(wd))
        assert torch.equal(wf, _bf(w))
        dx = torch.empty(n, 128, dtype=torch.bfloat16, device="cuda")
        gate._call("infx_gate_conv", gate._p(dy), n, seg, 128, gate._p(wd), kp, None, gate._p(dx), 1)
        _close(dx, xr.grad, 1.6e-2)
    cs = torch.zeros(128, device="cuda")
    gate._call("infx_gate_colsum", gate._p(dy), n, gate._p(cs))
    _close(cs, dy.float().sum(0), 1e-4)


def test_head_loss_relu_adam():
    gate = _gate()
    n = 5000
    g = torch.Generator().manual_seed(3)
    y = _bf(torch.relu(torch.randn(n, 128, generator=g))).cuda()
    hw = torch.randn(128, generator=g).cuda()
    hb = torch.randn(1, generator=g).cuda()
    s = torch.empty(n, device="cuda")
    gate._call("infx_gate_head", gate._p(y), n, gate._p(hw), gate._p(hb), gate._p(s))
    sref = y.float() @ hw + hb
    _close(s, sref, 1e-4)
    lab = (torch.rand(n, generator=g) < 0.2).float().cuda()
    pw = 3.0
    dy = torch.empty(n, 128, dtype=torch.bfloat16, device="cuda")
    dhw = torch.zeros(128, device="cuda")
    dhb = torch.zeros(1, device="cuda")
    loss = torch.zeros(1, device="cuda")
    gate._call("infx_gate_head_bwd", gate._p(s), gate._p(lab), gate._p(y), n, gate._p(hw), pw, gate._p(dy),
               gate._p(dhw), gate._p(dhb), gate._p(loss))
    sr = s.clone().requires_grad_(True)
    wgt = torch.where(lab > 0.5, torch.full_like(lab, pw), torch.ones_like(lab))
    lref = (wgt * torch.nn.functional.binary_cross_entropy_with_logits(sr, lab, reduction="none")).mean()
    lref.backward()
    assert abs(loss.item() - lref.item()) <= 1e-4 * max(1.0, lref.item())
    _close(dhw, y.float().t() @ sr.grad, 1e-3)
    assert abs(dhb.item() - sr.grad.sum().item()) <= 1e-4
    _close(dy, sr.grad[:, None] * hw[None, :], 1.6e-2)
    # ReLU backward
    g2 = dy.clone()
    gate._call("infx_gate_relu_bwd", gate._p(g2), gate._p(y), n * 128)
    assert torch.equal(g2, torch.where(y.float() > 0, dy, torch.zeros_like(dy)))
    # Adam
    w = torch.randn(1000, generator=g).cuda()
    gr = torch.randn(1000, generator=g).cuda()
    m = torch.zeros(1000, device="cuda")
    v = torch.zeros(1000, device="cuda")
    w0 = w.clone()
    gate._call("infx_gate_adam", gate._p(w), gate._p(gr), gate._p(m), gate._p(v), 1000, 0.01, 0.9, 0.999, 1e-8, 1)
    mref, vref = 0.1 * gr, 0.001 * gr * gr
    wref = w0 - 0.01 * (mref / 0.1) / (torch.sqrt(vref / 0.001) + 1e-8)
    _close(w, wref, 1e-5)


def test_training_learns_a_motif():
    """A planted 12-nt motif in random sequence: after 150 steps the gate
    ranks motif residues above the rest (AUC > 0.95)."""
    gate = _gate()
    rng = np.random.default_rng(0)
    motif = rng.integers(0, 4, 12).astype(np.uint8)
    net = gate.GateNet(seed=1)
    seg, nseg = 512, 16

    def batch():
        codes = rng.integers(0, 4, (nseg, seg)).astype(np.uint8)
        lab = np.zeros((nseg, seg), dtype=np.float32)
        for s in range(nseg):
            for _ in range(3):
                p = int(rng.integers(20, seg - 32))
                codes[s, p:p + 12] = motif
                lab[s, p:p + 12] = 1.0
        return torch.from_numpy(codes.ravel()).cuda(), torch.from_numpy(lab.ravel()).cuda(), lab.ravel()

    first = None
    for it in range(150):
        c, l, _ = batch()
        loss = net.train_step(c, l, seg, lr=2e-3, pos_weight=4.0)
        first = loss if first is None else first
    c, l, lab = batch()
    s, _ = net.forward(c, seg)
    s = s.cpu().numpy()
    pos, neg = s[lab > 0.5], s[lab < 0.5]
    auc = (pos[:, None] > neg[None, ::50]).mean()
    print("loss", first, "->", loss, "auc", auc)
    assert loss < 0.5 * first
    assert auc > 0.95
