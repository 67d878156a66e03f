"""Pins of the float64 warm-start network oracle (oracle/cnn_oracle.py):
the convolution against a direct sum over taps and channels, every gradient
against central finite differences, and the Adam step against its formula."""

import numpy as np

from oracle import cnn_oracle as cn


def _direct_conv(x, w, b):
    bsz, h, wd, ci = x.shape
    co = w.shape[2]
    out = np.zeros((bsz, h, wd, co))
    for bi in range(bsz):
        for i in range(h):
            for j in range(wd):
                acc = b.copy()
                for kh in range(3):
                    for kw in range(3):
                        ii, jj = i + kh - 1, j + kw - 1
                        if 0 <= ii < h and 0 <= jj < wd:
                            acc = acc + x[bi, ii, jj, :] @ w[3 * kh + kw]
                out[bi, i, j] = acc
    return out


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 4, 3))
    w = rng.standard_normal((9, 3, 2))
    b = rng.standard_normal(2)
    np.testing.assert_allclose(cn.conv3x3(x, w, b), _direct_conv(x, w, b), rtol=1e-12, atol=1e-12)


def _small_net(seed=1):
    rng = np.random.default_rng(seed)
    params = []
    for l in range(4):
        ci, co = cn.CHANNELS[l], cn.CHANNELS[l + 1]
        params.append((rng.standard_normal((9, ci, co)) * 0.2, rng.standard_normal(co) * 0.1))
    x = np.zeros((2, 5, 6, cn.CHANNELS[0]))
    x[..., :2] = rng.standard_normal((2, 5, 6, 2))
    target = rng.standard_normal((2, 5, 6))
    return params, x, target


def test_gradients_match_finite_differences():
    params, x, target = _small_net()
    loss, grads, _ = cn.loss_and_grads(x, target, params)
    rng = np.random.default_rng(7)
    h = 1e-6
    for l in range(4):
        w, b = params[l]
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            wp = w.copy(); wp[idx] += h
            wm = w.copy(); wm[idx] -= h
            pp = list(params); pp[l] = (wp, b)
            pm = list(params); pm[l] = (wm, b)
            fd = (cn.loss_and_grads(x, target, pp)[0] - cn.loss_and_grads(x, target, pm)[0]) / (2 * h)
            assert abs(fd - grads[l][0][idx]) <= 1e-6 * max(1.0, abs(fd))
        j = int(rng.integers(0, b.shape[0]))
        bp = b.copy(); bp[j] += h
        bm = b.copy(); bm[j] -= h
        pp = list(params); pp[l] = (w, bp)
        pm = list(params); pm[l] = (w, bm)
        fd = (cn.loss_and_grads(x, target, pp)[0] - cn.loss_and_grads(x, target, pm)[0]) / (2 * h)
        assert abs(fd - grads[l][1][j]) <= 1e-6 * max(1.0, abs(fd))


def test_input_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 4, 3))
    w = rng.standard_normal((9, 3, 2))
    b = rng.standard_normal(2)
    dz = rng.standard_normal((1, 4, 4, 2))
    dx, _, _ = cn.conv3x3_backward(x, w, dz)
    h = 1e-6
    for _ in range(8):
        idx = tuple(rng.integers(0, s) for s in x.shape)
        xp = x.copy(); xp[idx] += h
        xm = x.copy(); xm[idx] -= h
        fd = (np.sum(cn.conv3x3(xp, w, b) * dz) - np.sum(cn.conv3x3(xm, w, b) * dz)) / (2 * h)
        assert abs(fd - dx[idx]) <= 1e-6 * max(1.0, abs(fd))


def test_adam_first_step_is_signed_lr():
    params, x, target = _small_net(2)
    _, grads, _ = cn.loss_and_grads(x, target, params)
    state = [(np.zeros_like(w), np.zeros_like(w), np.zeros_like(b), np.zeros_like(b)) for w, b in params]
    new = cn.adam_step(params, grads, state, lr=1e-3, t=1)
    for (w0, _), (w1, _), (dw, _) in zip(params, new, grads):
        big = np.abs(dw) > 1e-3
        np.testing.assert_allclose((w0 - w1)[big], 1e-3 * np.sign(dw[big]), rtol=1e-4)


def test_padded_channels_stay_zero():
    params = cn.init_params(0)
    assert np.all(params[0][0][:, 2:, :] == 0.0)
    assert np.all(params[3][0][:, :, 1:] == 0.0)
    bp = np.ones((1, 4, 4), complex)
    g = np.ones((1, 4, 4))
    f = cn.features(bp, g)
    assert f.shape == (1, 4, 4, 16) and np.all(f[..., 2:] == 0)
