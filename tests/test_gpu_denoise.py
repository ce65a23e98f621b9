"""The denoiser's matrix-core kernels through the C ABI against the CPU
oracle (oracle/denoise_oracle.c, pinned by tests/test_denoise_oracle.py).

Single operations get the oracle's exact bf16 inputs, so the only difference
is fp32 accumulation order: |gpu - oracle| <= 1e-5 * sum |terms|.  The whole
network stores bf16 activations and gradients at the same points as the
oracle in rounding mode; the tolerance there allows a few bf16 ulps."""
import ctypes as C

import numpy as np
import pytest

from oracle import denoise as dno

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch(gpu):
    import torch
    return torch


def bf(torch, a):
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).cuda().contiguous()


def p(t):
    return C.c_void_p(t.data_ptr())


# upsampled inputs through the multi-slice kernels (maps of whole and partial
# 8 x 32 tiles) and through the one-slice one
@pytest.mark.parametrize("up,c1,c2,cout,relu,hw", [(0, 32, 0, 32, 1, (16, 24)), (0, 64, 0, 64, 0, (8, 8)),
                                                  (1, 64, 32, 32, 1, (16, 16)), (1, 128, 64, 64, 1, (8, 16)),
                                                  (1, 64, 32, 64, 0, (12, 40)), (1, 32, 0, 32, 1, (8, 64)),
                                                  (0, 32, 0, 128, 1, (6, 10))])
def test_conv3x3_matches_oracle(torch, up, c1, c2, cout, relu, hw):
    from fornos_amd._abi import check, lib
    rng = np.random.default_rng(c1 + cout)
    n, (h, w) = 2, hw
    ah, aw = (h // 2, w // 2) if up else (h, w)
    a = bf(torch, rng.standard_normal((n, ah, aw, c1)))
    b = bf(torch, rng.standard_normal((n, h, w, c2))) if c2 else None
    wt = bf(torch, rng.standard_normal((cout, 9, c1 + c2)) * 0.1)
    bias = torch.as_tensor(rng.standard_normal(cout).astype(np.float32)).cuda()
    out = torch.empty((n, h, w, cout), dtype=torch.float32, device="cuda")
    check(lib().fornos_dn_conv3x3(p(a), c1, up, p(b) if b is not None else None, c2, n, h, w, p(wt), p(bias),
                                  cout, relu, p(out)))
    f = lambda t: t.float().cpu().numpy().astype(np.float64)
    ref = dno.conv3x3(f(a), c1, up, f(b) if b is not None else None, c2, n, h, w, f(wt),
                      bias.cpu().numpy().astype(np.float64), cout, relu)
    mag = dno.conv3x3(np.abs(f(a)), c1, up, np.abs(f(b)) if b is not None else None, c2, n, h, w,
                      np.abs(f(wt)), np.abs(bias.cpu().numpy()).astype(np.float64), cout, 0)
    err = np.abs(out.cpu().numpy() - ref)
    assert np.all(err <= 1e-5 * mag + 1e-6), float(np.max(err / (mag + 1e-6)))


@pytest.mark.parametrize("up,c1,c2,cout,hw", [(0, 32, 0, 32, (16, 16)), (1, 64, 32, 32, (8, 12)),
                                              (0, 128, 0, 64, (32, 32)), (1, 256, 128, 128, (4, 4)),
                                              (0, 32, 0, 32, (96, 128))])
def test_wgrad_matches_oracle(torch, up, c1, c2, cout, hw):
    from fornos_amd._abi import check, lib
    rng = np.random.default_rng(c1 + c2 + cout)
    n, (h, w) = 3, hw
    ah, aw = (h // 2, w // 2) if up else (h, w)
    a = bf(torch, rng.standard_normal((n, ah, aw, c1)))
    b = bf(torch, rng.standard_normal((n, h, w, c2))) if c2 else None
    g = bf(torch, rng.standard_normal((n, h, w, cout)))
    dw = torch.empty((cout, 9, c1 + c2), dtype=torch.float32, device="cuda")
    db = torch.empty(cout, dtype=torch.float32, device="cuda")
    check(lib().fornos_dn_conv3x3_wgrad(p(a), c1, up, p(b) if b is not None else None, c2, n, h, w, p(g), cout,
                                        p(dw), p(db)))
    f = lambda t: t.float().cpu().numpy().astype(np.float64)
    rw, rb = dno.conv3x3_wgrad(f(a), c1, up, f(b) if b is not None else None, c2, n, h, w, f(g), cout)
    mw, mb = dno.conv3x3_wgrad(np.abs(f(a)), c1, up, np.abs(f(b)) if b is not None else None, c2, n, h, w,
                               np.abs(f(g)), cout)
    assert np.all(np.abs(dw.cpu().numpy() - rw) <= 1e-5 * mw + 1e-6)
    assert np.all(np.abs(db.cpu().numpy() - rb) <= 1e-5 * mb + 1e-6)
    # the splits of K are added in a fixed order: a second run gives the same bits
    dw2, db2 = torch.empty_like(dw), torch.empty_like(db)
    check(lib().fornos_dn_conv3x3_wgrad(p(a), c1, up, p(b) if b is not None else None, c2, n, h, w, p(g), cout,
                                        p(dw2), p(db2)))
    assert torch.equal(dw.view(torch.int32), dw2.view(torch.int32))
    assert torch.equal(db.view(torch.int32), db2.view(torch.int32))


def small_batch(n, h, w, seed):
    rng = np.random.default_rng(seed)
    inp = rng.standard_normal((n, h, w, dno.IN)).astype(np.float32) * 0.5
    inp[..., dno.MASK] = (rng.random((n, h, w)) < 0.8)
    tgt = (rng.standard_normal((n, h, w, dno.OUT)) * 0.5).astype(np.float32)
    return inp, tgt


def test_network_forward_and_gradient_match_oracle(torch):
    from fornos_amd.denoise import Denoiser
    levels, base, n, h, w = 3, 32, 2, 16, 16
    net = Denoiser(levels, base, seed=5)
    params = net.get_params()
    # non-trivial biases and head so every path carries signal
    rng = np.random.default_rng(1)
    for L in net.layers():
        cin = L["c1"] + L["c2"]
        o = L["offset"] + L["cout"] * 9 * cin
        params[o:o + L["cout"]] = rng.standard_normal(L["cout"]) * 0.05
    params[-4 * base - 4:] = rng.standard_normal(4 * base + 4) * 0.2
    net.set_params(params)
    inp, tgt = small_batch(n, h, w, 2)
    ti, tt = torch.as_tensor(inp).cuda(), torch.as_tensor(tgt).cuda()
    out, _ = net.forward(ti)
    ref = dno.forward(levels, base, params.astype(np.float64), inp.astype(np.float64), round=True)
    o = out.cpu().numpy()
    assert np.max(np.abs(o - ref)) <= 2e-2 * np.max(np.abs(ref))
    assert np.mean(np.abs(o - ref)) <= 2e-3 * np.mean(np.abs(ref))
    loss = net.backward(ti, tt)
    g = net.grads()
    rl, rg, _ = dno.backward(levels, base, params.astype(np.float64), inp.astype(np.float64),
                             tgt.astype(np.float64), round=True)
    assert abs(loss - rl) <= 1e-3 * abs(rl)
    for L in net.layers():
        cin = L["c1"] + L["c2"]
        sl = slice(L["offset"], L["offset"] + L["cout"] * 9 * cin)
        a, b = g[sl].astype(np.float64), rg[sl]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        assert cos > 0.995 and rel < 0.1, (L, cos, rel)
    hs = slice(len(g) - 4 * base - 4, len(g))
    assert np.allclose(g[hs], rg[hs], rtol=2e-2, atol=1e-4 * np.max(np.abs(rg[hs])))


def test_training_reduces_loss(torch):
    from fornos_amd.denoise import Denoiser
    net = Denoiser(4, 32, seed=3)
    inp, _ = small_batch(4, 32, 32, 9)
    # target: a smooth function of the inputs the head alone cannot produce
    tgt = np.zeros(inp.shape[:3] + (4,), np.float32)
    tgt[..., 0] = np.tanh(inp[..., 4] + inp[..., 5])
    tgt[..., 1:] = inp[..., 7:10] * 0.5
    ti, tt = torch.as_tensor(inp).cuda(), torch.as_tensor(tgt).cuda()
    losses = []
    for it in range(60):
        losses.append(net.backward(ti, tt))
        net.adam(2e-3)
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-5:]) < 0.6 * np.mean(losses[:3]), losses[::10]
