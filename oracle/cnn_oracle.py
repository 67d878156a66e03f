"""Float64 restatement of the learned warm-start network (test oracle).

TEST INFRASTRUCTURE ONLY.  The reference has no learned component; this
network is specified by the task (4 conv layers, 3x3 kernels, 32 channels,
inputs |A^H y| and log gamma, output a per-cell log-variance correction).
The oracle states, in plain NumPy, the exact arithmetic the HIP kernels
implement so their forward outputs and gradients can be checked:

  features  f0 = log(|bp|^2 / mean|bp|^2 + 1e-6), f1 = log(gamma / mean gamma)
            (both clipped to [-30, 30]); channels 2..15 are zero padding
  layer l   z_l = conv3x3(h_{l-1}, W_l) + b_l  ("same" zero padding)
            h_l = relu(z_l) for l < 4, h_4 = z_4 (channel 0 is the output)
  loss      L = mean over cells of (h_4[..., 0] - target)^2

Layouts: activations [B, H, W, C] (NHWC); weights W[tap, ci, co] with
tap = 3 * kh + kw, input cell (h + kh - 1, w + kw - 1).

Pins: tests/test_oracle_cnn.py checks every gradient against central finite
differences and the convolution against a direct four-loop sum.
"""

from __future__ import annotations

import numpy as np

CHANNELS = (16, 32, 32, 32, 16)   # padded input / hidden / padded output
FEATURE_CLIP = 30.0


def features(bp: np.ndarray, gamma: np.ndarray) -> np.ndarray:
    """[B, H, W] complex backprojection and positive gamma -> [B, H, W, 16]."""
    e = np.abs(bp) ** 2
    f0 = np.log(e / np.mean(e, axis=(1, 2), keepdims=True) + 1e-6)
    f1 = np.log(gamma / np.mean(gamma, axis=(1, 2), keepdims=True))
    out = np.zeros(bp.shape + (CHANNELS[0],))
    out[..., 0] = np.clip(f0, -FEATURE_CLIP, FEATURE_CLIP)
    out[..., 1] = np.clip(f1, -FEATURE_CLIP, FEATURE_CLIP)
    return out


def conv3x3(x: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x [B, H, W, Ci], w [9, Ci, Co], b [Co] -> [B, H, W, Co]."""
    bsz, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.broadcast_to(b, (bsz, h, wd, w.shape[2])).copy()
    for kh in range(3):
        for kw in range(3):
            out += xp[:, kh:kh + h, kw:kw + wd, :] @ w[3 * kh + kw]
    return out


def conv3x3_backward(x: np.ndarray, w: np.ndarray, dz: np.ndarray):
    """Gradients of conv3x3 w.r.t. x, w and b given dL/dz."""
    bsz, h, wd, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for kh in range(3):
        for kw in range(3):
            tap = 3 * kh + kw
            patch = xp[:, kh:kh + h, kw:kw + wd, :]
            dw[tap] = np.einsum("bhwi,bhwo->io", patch, dz)
            dxp[:, kh:kh + h, kw:kw + wd, :] += dz @ w[tap].T
    db = dz.sum(axis=(0, 1, 2))
    return dxp[:, 1:-1, 1:-1, :], dw, db


def init_params(seed: int = 0):
    """He-normal weights (zero on the padded input channels), zero bias."""
    rng = np.random.default_rng(seed)
    params = []
    fan_in_ch = (2, 32, 32, 32)
    for l in range(4):
        ci, co = CHANNELS[l], CHANNELS[l + 1]
        w = np.zeros((9, ci, co))
        real_ci = fan_in_ch[l]
        real_co = 1 if l == 3 else co
        w[:, :real_ci, :real_co] = rng.standard_normal((9, real_ci, real_co)) * np.sqrt(2.0 / (9 * real_ci))
        if l == 3:
            w *= 0.1   # start near the identity warm start (zero correction)
        params.append((w, np.zeros(co)))
    return params


def forward(x: np.ndarray, params):
    """Returns (output channel 0 [B, H, W], cache of layer inputs/pre-acts)."""
    cache = []
    h = x
    for l, (w, b) in enumerate(params):
        z = conv3x3(h, w, b)
        cache.append((h, z))
        h = np.maximum(z, 0.0) if l < 3 else z
    return h[..., 0], cache


def loss_and_grads(x: np.ndarray, target: np.ndarray, params):
    """Mean squared error on channel 0 and its gradients."""
    out, cache = forward(x, params)
    diff = out - target
    loss = float(np.mean(diff ** 2))
    dh = np.zeros(out.shape + (CHANNELS[4],))
    dh[..., 0] = 2.0 * diff / diff.size
    grads = [None] * 4
    for l in range(3, -1, -1):
        h_in, z = cache[l]
        dz = dh if l == 3 else dh * (z > 0.0)
        dx, dw, db = conv3x3_backward(h_in, params[l][0], dz)
        grads[l] = (dw, db)
        dh = dx
    return loss, grads, out


def adam_step(params, grads, state, lr, t, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), in place on `state` [(m_w, v_w, m_b, v_b)]."""
    new = []
    c1 = 1.0 - beta1 ** t
    c2 = 1.0 - beta2 ** t
    for l, ((w, b), (dw, db)) in enumerate(zip(params, grads)):
        mw, vw, mb, vb = state[l]
        mw[:] = beta1 * mw + (1 - beta1) * dw
        vw[:] = beta2 * vw + (1 - beta2) * dw * dw
        mb[:] = beta1 * mb + (1 - beta1) * db
        vb[:] = beta2 * vb + (1 - beta2) * db * db
        w = w - lr * (mw / c1) / (np.sqrt(vw / c2) + eps)
        b = b - lr * (mb / c1) / (np.sqrt(vb / c2) + eps)
        new.append((w, b))
    return new
