"""Pins of the denoiser oracle (oracle/denoise_oracle.c): the convolution and
its weight gradient against torch's CPU conv2d and autograd, the whole network
against a torch restatement built from library layers (conv2d, avg_pool2d,
nearest interpolation, cat), and its gradient against central finite
differences.  CPU only."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import denoise as dno

torch.set_default_dtype(torch.float64)


def t_conv(x_nhwc, w, bias, relu):
    # w: [cout][9][cin] -> [cout][cin][3][3]
    cout, _, cin = w.shape
    wt = torch.as_tensor(w).reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
    y = F.conv2d(x_nhwc.permute(0, 3, 1, 2), wt, None if bias is None else torch.as_tensor(bias), padding=1)
    y = y.permute(0, 2, 3, 1)
    return F.relu(y) if relu else y


def up2(x):
    return F.interpolate(x.permute(0, 3, 1, 2), scale_factor=2, mode="nearest").permute(0, 2, 3, 1)


@pytest.mark.parametrize("up", [0, 1])
def test_conv_and_wgrad_match_torch(up):
    rng = np.random.default_rng(up)
    n, h, w, c1, c2, cout = 2, 6, 8, 32, 32, 32
    a = rng.standard_normal((n, h // 2 if up else h, w // 2 if up else w, c1))
    b = rng.standard_normal((n, h, w, c2))
    wts = rng.standard_normal((cout, 9, c1 + c2)) * 0.1
    bias = rng.standard_normal(cout)
    out = dno.conv3x3(a, c1, up, b, c2, n, h, w, wts, bias, cout, True)
    ta = torch.as_tensor(a)
    x = torch.cat([up2(ta) if up else ta, torch.as_tensor(b)], dim=3)
    ref = t_conv(x, wts, bias, True)
    assert np.allclose(out, ref.numpy(), rtol=1e-12, atol=1e-12)
    g = rng.standard_normal((n, h, w, cout))
    dw, db = dno.conv3x3_wgrad(a, c1, up, b, c2, n, h, w, g, cout)
    W = torch.as_tensor(wts).clone().requires_grad_(True)
    B = torch.as_tensor(bias).clone().requires_grad_(True)
    y = t_conv(x, W, B, False)
    (y * torch.as_tensor(g)).sum().backward()
    assert np.allclose(dw, W.grad.numpy(), rtol=1e-10, atol=1e-10)
    assert np.allclose(db, B.grad.numpy(), rtol=1e-10, atol=1e-10)


def layers(levels, base):
    out, off = [], 0
    def add(c1, c2, co, upf):
        nonlocal off
        L = {"c1": c1, "c2": c2, "cout": co, "up": upf, "w": off}
        off += co * 9 * (c1 + c2)
        L["b"] = off
        off += co
        out.append(L)
    for l in range(levels):
        add(32 if l == 0 else base << (l - 1), 0, base << l, 0)
        add(base << l, 0, base << l, 0)
    for l in range(levels - 2, -1, -1):
        add(base << (l + 1), base << l, base << l, 1)
        add(base << l, 0, base << l, 0)
    return out, off


def torch_unet(levels, base, params, inp):
    """The network from library layers, NHWC tensors."""
    Ls, off = layers(levels, base)
    p = params
    def conv(i, x):
        L = Ls[i]
        cin = L["c1"] + L["c2"]
        wt = p[L["w"]:L["w"] + L["cout"] * 9 * cin].reshape(L["cout"], 9, cin)
        return t_conv(x, wt, p[L["b"]:L["b"] + L["cout"]], True)
    x = F.pad(inp, (0, 32 - inp.shape[3]))
    E = []
    for l in range(levels):
        if l:
            x = F.avg_pool2d(E[-1].permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)
        e0 = conv(2 * l, x)
        E.append(conv(2 * l + 1, e0))
    below = E[-1]
    i = 2 * levels
    for l in range(levels - 2, -1, -1):
        d0 = conv(i, torch.cat([up2(below), E[l]], dim=3))
        below = conv(i + 1, d0)
        i += 2
    hw = p[off:off + 4 * base].reshape(4, base)
    hb = p[off + 4 * base:off + 4 * base + 4]
    return inp[..., :4] + hb + below @ hw.T


def sample(levels, n, h, w, seed):
    rng = np.random.default_rng(seed)
    inp = rng.standard_normal((n, h, w, dno.IN)) * 0.5
    inp[..., dno.MASK] = (rng.random((n, h, w)) < 0.8).astype(np.float64)
    tgt = rng.standard_normal((n, h, w, dno.OUT)) * 0.5
    return inp, tgt


def init_params(levels, base, seed):
    Ls, total = layers(levels, base)
    rng = np.random.default_rng(seed)
    p = np.zeros(total + 4 * base + 4)
    for i, L in enumerate(Ls):
        cin = L["c1"] + L["c2"]
        p[L["w"]:L["w"] + L["cout"] * 9 * cin] = rng.standard_normal(L["cout"] * 9 * cin) * np.sqrt(2.0 / (9 * cin))
        p[L["b"]:L["b"] + L["cout"]] = rng.standard_normal(L["cout"]) * 0.05
    p[-4 * base - 4:] = rng.standard_normal(4 * base + 4) * 0.1
    return p


@pytest.mark.parametrize("levels", [2, 3])
def test_network_and_gradient_match_torch(levels):
    base, n, h, w = 32, 2, 8, 8
    assert dno.param_count(levels, base) == layers(levels, base)[1] + 4 * base + 4
    params = init_params(levels, base, 7)
    inp, tgt = sample(levels, n, h, w, 3)
    out = dno.forward(levels, base, params, inp, round=False)
    P = torch.as_tensor(params).clone().requires_grad_(True)
    ref = torch_unet(levels, base, P, torch.as_tensor(inp))
    assert np.allclose(out, ref.detach().numpy(), rtol=1e-10, atol=1e-10)
    mask = torch.as_tensor(inp[..., dno.MASK] > 0.5)
    loss_t = ((ref - torch.as_tensor(tgt)).abs().sum(-1) * mask).sum() / mask.sum()
    loss_t.backward()
    loss, g, _ = dno.backward(levels, base, params, inp, tgt, round=False)
    assert abs(loss - loss_t.item()) < 1e-10
    assert np.allclose(g, P.grad.numpy(), rtol=1e-8, atol=1e-10)


def test_gradient_matches_finite_differences():
    levels, base, n, h, w = 2, 32, 1, 4, 4
    params = init_params(levels, base, 11)
    inp, tgt = sample(levels, n, h, w, 5)
    loss, g, _ = dno.backward(levels, base, params, inp, tgt, round=False)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(len(params), 40, replace=False), np.arange(len(params) - 8, len(params))])
    eps = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _, _ = dno.backward(levels, base, pp, inp, tgt, round=False)
        lm, _, _ = dno.backward(levels, base, pm, inp, tgt, round=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_rounding_mode_rounds_activations_only_slightly():
    levels, base = 3, 32
    params = init_params(levels, base, 1)
    inp, tgt = sample(levels, 1, 8, 8, 2)
    a = dno.forward(levels, base, params, inp, round=False)
    b = dno.forward(levels, base, params, inp, round=True)
    assert not np.array_equal(a, b)
    assert np.max(np.abs(a - b)) < 0.05 * np.max(np.abs(a))
    assert np.array_equal(dno.bf16(np.array([1.0, 1.00390625, 1.005859375])), [1.0, 1.0, 1.0078125])
