"""Warm-start network kernels (BF16 MFMA convolutions, backward, Adam)
against the float64 oracle (oracle/cnn_oracle.py).

Two comparisons per quantity: against the oracle evaluated on the same
bf16-rounded tensors the kernels see (weights, inputs, hidden activations and
backward signals rounded to bf16 exactly where the kernels round), which pins
the kernel arithmetic to FP32-accumulation differences (relative 2e-3); and
against the plain float64 oracle, which bounds the bf16 error of the whole
network (relative 5e-2).
"""

import numpy as np
import pytest
import torch

from oracle import cnn_oracle as cn

pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def to_bits(a):
    u = np.asarray(bf16(a), np.float32).view(np.uint32)
    return torch.from_numpy((u >> 16).astype(np.uint16).view(np.int16).copy())


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-30))


def _setup(device, b=2, h=40, w=80, seed=0):
    from torch_sbl import learned
    rng = np.random.default_rng(seed)
    p = learned.init_params(seed)
    layers = learned.unflatten(p)
    # non-zero biases and a stronger output layer so every path matters
    flat = []
    for l, (wl, bl) in enumerate(layers):
        wl = wl.copy()
        if l == 3:
            wl *= 10.0
        bl = np.zeros_like(bl)
        co = 1 if l == 3 else bl.size
        bl[:co] = rng.standard_normal(co) * 0.05
        flat += [wl.reshape(-1), bl]
    p = np.concatenate(flat).astype(np.float32)
    net = learned.WarmStartNet(device, p)
    x = np.zeros((b, h, w, 16))
    x[..., :2] = rng.standard_normal((b, h, w, 2))
    target = rng.standard_normal((b, h, w)).astype(np.float32)
    return net, learned.unflatten(p), bf16(x), target


def _emulated_forward(x, params):
    """Forward on the kernels' bf16 roundings: weights and hidden outputs."""
    hs, zs = [x], []
    h = x
    for l, (w, b) in enumerate(params):
        z = cn.conv3x3(h, bf16(w), b)
        zs.append(z)
        h = bf16(np.maximum(z, 0.0)) if l < 3 else z
        hs.append(h)
    return hs, zs


@pytest.mark.parametrize("shape", [(2, 40, 80), (1, 64, 64), (1, 17, 33)])
def test_forward_matches_oracle(device, shape):
    b, h, w = shape
    print("shape", shape, flush=True)
    net, params, x, _ = _setup(device, b, h, w)
    out = net.forward(to_bits(x).to(device)).cpu().numpy()
    hs, _ = _emulated_forward(x, params)
    assert _rel(out, hs[-1][..., 0]) < 2e-3
    ref, _ = cn.forward(x, [(np.float64(wl), np.float64(bl)) for wl, bl in params])
    assert _rel(out, ref) < 5e-2


def test_gradients_match_oracle(device):
    net, params, x, target = _setup(device, 2, 32, 64, seed=3)
    loss = float(net.loss_and_grads(to_bits(x).to(device), torch.from_numpy(target).to(device)).item())
    g = net.grads.cpu().numpy()
    from torch_sbl import learned
    got = learned.unflatten(g)
    # emulation of the kernels' roundings through the backward pass
    hs, zs = _emulated_forward(x, params)
    diff = hs[-1][..., 0] - target
    assert loss == pytest.approx(float(np.mean(diff ** 2)), rel=1e-4)
    dz = np.zeros(diff.shape + (16,))
    dz[..., 0] = bf16(2.0 * diff / diff.size)
    for l in range(3, -1, -1):
        _, dw, db = cn.conv3x3_backward(hs[l], bf16(params[l][0]), dz)
        assert _rel(got[l][0], dw) < 2e-3, l
        assert _rel(got[l][1], db) < 2e-3, l
        if l > 0:
            dx, _, _ = cn.conv3x3_backward(hs[l], bf16(params[l][0]), dz)
            dz = bf16(dx * (hs[l] > 0))
    # against the float64 oracle (bf16 error of the whole backward pass)
    _, grads, _ = cn.loss_and_grads(x, np.float64(target), [(np.float64(wl), np.float64(bl)) for wl, bl in params])
    for l in range(4):
        assert _rel(got[l][0], grads[l][0]) < 5e-2, l


def test_adam_matches_oracle(device):
    net, params, x, target = _setup(device, 1, 16, 64, seed=5)
    net.train_step(to_bits(x).to(device), torch.from_numpy(target).to(device), lr=1e-3)
    g = net.grads.cpu().numpy()
    from torch_sbl import learned
    state = [(np.zeros_like(w, np.float64), np.zeros_like(w, np.float64), np.zeros_like(b, np.float64),
              np.zeros_like(b, np.float64)) for w, b in params]
    new = cn.adam_step([(np.float64(w), np.float64(b)) for w, b in params],
                       [(np.float64(a), np.float64(c)) for a, c in learned.unflatten(g)], state, 1e-3, 1)
    got = learned.unflatten(net.params.cpu().numpy())
    for (w, b), (gw, gb) in zip(new, got):
        np.testing.assert_allclose(gw, w, rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(gb, b, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(device):
    """A learnable target (a local function of the input features) is fit."""
    from torch_sbl import learned
    rng = np.random.default_rng(9)
    net = learned.WarmStartNet(device, seed=1)
    x = np.zeros((4, 32, 64, 16))
    x[..., :2] = rng.standard_normal((4, 32, 64, 2))
    x = bf16(x)
    target = (0.5 * x[..., 0] - 0.3 * np.roll(x[..., 1], 1, axis=2)).astype(np.float32)
    xb = to_bits(x).to(device)
    tt = torch.from_numpy(target).to(device)
    first = float(net.train_step(xb, tt, 3e-3).item())
    for _ in range(150):
        last = float(net.train_step(xb, tt, 3e-3).item())
    print("loss", first, "->", last, flush=True)
    assert last < 0.1 * first
