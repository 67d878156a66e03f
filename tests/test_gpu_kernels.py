"""The five Triton-wrapper equivalents (torch_sbl.hip_kernels) against the
oracle's float64 restatement of triton_kernels.py:36-417."""

import numpy as np
import pytest
import torch

from oracle import fbme_oracle as fo

pytestmark = pytest.mark.gpu


def _c(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


@pytest.mark.parametrize("b,m", [(1, 576), (3, 25600), (2, 1)])
def test_message_update(device, b, m):
    from torch_sbl import hip_kernels as hk
    rng = np.random.default_rng(0)
    y, p, old = _c(rng, b, m), _c(rng, b, m), _c(rng, b, m)
    dv = rng.uniform(0.5, 2, b).astype(np.float32)
    nz = rng.uniform(0.1, 1, b).astype(np.float32)
    msg, tau = hk.uamp_message_update(_t(y, device), _t(p, device), _t(old, device),
                                      _t(dv, device), _t(nz, device), damping=0.05)
    st = dict(prediction=p.astype(complex), message=old.astype(complex),
              data_variance=dv.astype(float))
    prec = 1 / (dv + nz)
    ref = 0.05 * old + 0.95 * prec[:, None] * (y - (p - dv[:, None] * old))
    np.testing.assert_allclose(msg.cpu().numpy(), ref, rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(tau.cpu().numpy(), (dv + nz) / m, rtol=1e-6)


@pytest.mark.parametrize("b,n,fused", [(2, 65536, True), (1, 1024, False), (3, 300, True)])
def test_denoise_and_reductions(device, b, n, fused):
    from torch_sbl import hip_kernels as hk
    rng = np.random.default_rng(1)
    gam = rng.uniform(1e-3, 2, (b, n)).astype(np.float32)
    tau = rng.uniform(0.01, 0.5, b).astype(np.float32)
    lifted, old_mean = _c(rng, b, n), _c(rng, b, n)
    old_var = rng.uniform(1e-3, 1, (b, n)).astype(np.float32)
    floor = np.full(b, 1e-6, np.float32)
    out = hk.uamp_denoise(_t(gam, device), _t(tau, device), _t(lifted, device),
                          _t(old_mean, device), _t(old_var, device), damping=0.05,
                          tiny=fo.F32_TINY, prior_floor=_t(floor, device) if fused else None,
                          hyper_shape=0.0, hyper_rate=0.0, prior_damping=0.15,
                          prior_maximum=1e12)
    state = dict(mean=old_mean.astype(complex), variance=old_var.astype(float),
                 message=np.zeros((b, 1)), prediction=np.zeros((b, 1)),
                 data_variance=np.ones(b))
    # oracle denoise with the same tau: feed tau via data_variance/noise
    g64 = gam.astype(float)
    t64 = tau.astype(float)
    denom = np.maximum(g64 + t64[:, None], fo.F32_TINY)
    cm = g64 / denom * (old_mean + t64[:, None] * lifted)
    cv = g64 * t64[:, None] / denom
    mu = 0.05 * old_mean + 0.95 * cm
    v = np.maximum(0.05 * old_var + 0.95 * cv, fo.F32_TINY)
    rel = np.clip(1 - v / g64, 0, 1)
    np.testing.assert_allclose(out[0].cpu().numpy(), mu, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(out[1].cpu().numpy(), v, rtol=1e-5)
    np.testing.assert_allclose(out[2].cpu().numpy(), rel.sum(1), rtol=1e-5)
    np.testing.assert_allclose(out[3].cpu().numpy(), v.sum(1), rtol=1e-5)
    if fused:
        cp = np.maximum(1e-6, np.minimum(1e12, np.abs(mu) ** 2 + v))
        ref = np.maximum(0.15 * g64 + 0.85 * cp, 1e-6)
        np.testing.assert_allclose(out[4].cpu().numpy(), ref, rtol=1e-5)
    else:
        assert out[4] is None
    del state


def test_prior_update(device):
    from torch_sbl import hip_kernels as hk
    rng = np.random.default_rng(2)
    b, n = 2, 5000
    mu = _c(rng, b, n)
    pv = rng.uniform(0, 1, (b, n)).astype(np.float32)
    old = rng.uniform(0, 1, (b, n)).astype(np.float32)
    floor = np.array([1e-3, 0.5], np.float32)
    out = hk.uamp_prior_update(_t(mu, device), _t(pv, device), _t(old, device),
                               _t(floor, device), hyper_shape=1e-8, hyper_rate=1e-8,
                               damping=0.15, maximum=1e12).cpu().numpy()
    cand = (np.abs(mu.astype(complex)) ** 2 + pv + 1e-8) / (1 + 1e-8)
    cand = np.maximum(floor[:, None], np.minimum(1e12, cand))
    ref = np.maximum(0.15 * old + 0.85 * cand, floor[:, None])
    np.testing.assert_allclose(out, ref, rtol=2e-6)


@pytest.mark.parametrize("evidence", [False, True])
def test_residual_and_noise_update(device, evidence):
    from torch_sbl import hip_kernels as hk
    rng = np.random.default_rng(3)
    b, m = 3, 25600
    y, p = _c(rng, b, m), _c(rng, b, m)
    e = hk.complex_residual_energy(_t(y, device), _t(p, device)).cpu().numpy()
    ref_e = np.sum(np.abs(y.astype(complex) - p) ** 2, axis=1)
    np.testing.assert_allclose(e, ref_e, rtol=1e-6)
    old = rng.uniform(0.1, 1, b).astype(np.float32)
    dof = rng.uniform(1, 100, b).astype(np.float32)
    floor = np.full(b, 1e-6, np.float32)
    out = hk.uamp_noise_update(_t(y, device), _t(p, device), _t(old, device),
                               _t(dof, device), _t(floor, device), noise_shape=0.0,
                               noise_rate=0.0, damping=0.15, evidence_update=evidence,
                               epsilon=fo.F32_EPS).cpu().numpy()
    if evidence:
        cand = ref_e / np.maximum(m - dof, fo.F32_EPS)
    else:
        cand = (ref_e + old * dof) / m
    ref = 0.15 * old + 0.85 * np.maximum(cand, floor)
    np.testing.assert_allclose(out, ref, rtol=2e-6)


def test_contract_errors(device):
    from torch_sbl import hip_kernels as hk
    y = torch.zeros(2, 8, dtype=torch.complex64, device=device)
    with pytest.raises(ValueError):
        hk.uamp_message_update(y, y, y[:, :4].contiguous(), torch.ones(2, device=device),
                               torch.ones(2, device=device), damping=0.1)
    with pytest.raises(ValueError):
        hk.complex_residual_energy(y, y.to(torch.complex128))
