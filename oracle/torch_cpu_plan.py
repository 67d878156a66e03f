"""FBME-Fast plan on the host CPU in PyTorch (complex64 sweep, batched over
scenes, every host thread): the CPU baseline the benchmark times beside the
GPU plan.

TEST INFRASTRUCTURE ONLY (the benchmark's cpu_baseline leg and tests/): the
same algorithm as oracle/fbme_oracle.py (uamp_sweep + polish with the Fast
settings), which follows torch_sbl/solver.py:1063-1925 (init, UAMP loop,
final E-step; Torch UAMP branch 2808-2850) and 3961-4511 (zero-tail
active-set polish, evidence gate, support-fixed EM), restated on torch's
CPU kernels: A x = fft2(x)[:M1, :M2], A^H v = N ifft2(pad(v)) (operators.py:
237-280), float32 elementwise UAMP algebra, FP64 reduced posteriors on the
32 polished atoms with the closed-form Dirichlet Gram.  Pinned against the
float64 oracle in tests/test_oracle_golden.py.
"""

from __future__ import annotations

from math import ceil

import numpy as np
import torch

F32_EPS = float(np.finfo(np.float32).eps)
F32_TINY = float(np.finfo(np.float32).tiny)


def _dirichlet(m: int, n: int) -> torch.Tensor:
    """D[d] = sum_{j<m} exp(2 pi i j d / n), d = 0..n-1, float64 (phases
    reduced mod n)."""
    d = np.arange(n)[:, None]
    j = np.arange(m)[None, :]
    return torch.from_numpy(np.exp(2j * np.pi * ((d * j) % n) / n).sum(1))


def fast_plan(y: torch.Tensor, grid, *, polish_size: int = 32, uamp_damping: float = 0.05,
              damping: float = 0.15, min_variance: float = 1e-12, max_precision: float = 1e12,
              blend: float = 0.1, noise_fraction: float = 1e-2, scale: float = 0.8, min_iters: int = 3,
              evaluations: int = 2, margin_factor: float = 64.0, jitter: float = 8.0) -> dict:
    """y: [B, M1, M2] complex64 (CPU).  Returns x [B, N1, N2], noise [B]."""
    eps, tiny = F32_EPS, F32_TINY
    b = y.shape[0]
    m1, m2 = y.shape[-2:]
    n1, n2 = grid
    n, m = n1 * n2, m1 * m2
    y = y.reshape(b, m1, m2).to(torch.complex64)
    yf = y.reshape(b, m)

    def fwd(x):
        return torch.fft.fft2(x.reshape(b, n1, n2))[:, :m1, :m2].reshape(b, m)

    def adj(v):
        return (torch.fft.ifft2(v.reshape(b, m1, m2), s=(n1, n2)) * n).reshape(b, n)

    # init (solver.py:1063-1189)
    power = (yf.abs() ** 2).mean(1).clamp_min(eps)
    floor = (power * min_variance).clamp_min(tiny)
    bp = adj(yf)
    energy = bp.abs() ** 2
    gamma = blend * energy / energy.sum(1, keepdim=True).clamp_min(eps) * power[:, None] \
        + (1.0 - blend) * power[:, None] / n
    gamma = torch.maximum(gamma, floor[:, None])
    noise = torch.maximum(noise_fraction * power, floor)
    mean = torch.zeros_like(bp)
    var = gamma.clone()
    msg = torch.zeros_like(yf)
    pred = torch.zeros_like(yf)
    dv = gamma.sum(1)
    t_u = max(min_iters, ceil(scale * n / m))
    max_dof = m * (1.0 - 10.0 * eps)

    def step(it, gamma, noise, ard):
        nonlocal mean, var, msg, pred, dv
        v_old = var.clamp_min(tiny)
        d = dv.clamp_min(tiny)
        total = d + noise
        cand = (yf - (pred - d[:, None] * msg)) / total[:, None]
        msg = uamp_damping * msg + (1.0 - uamp_damping) * cand
        tau = (total / m).clamp_min(tiny)
        lifted = bp * ((1.0 - uamp_damping) / total)[:, None] if it == 1 else adj(msg)
        den = (gamma + tau[:, None]).clamp_min(tiny)
        cm = (gamma / den) * (mean + tau[:, None] * lifted)
        cv = gamma * tau[:, None] / den
        mean = uamp_damping * mean + (1.0 - uamp_damping) * cm
        var = (uamp_damping * v_old + (1.0 - uamp_damping) * cv).clamp_min(tiny)
        dof = (1.0 - var / gamma.clamp_min(tiny)).clamp(0.0, 1.0).sum(1)
        dv = var.sum(1)
        nxt = None
        if ard:
            cp = (mean.abs() ** 2 + var).clamp(max=max_precision)
            cp = torch.maximum(cp, floor[:, None])
            nxt = torch.maximum(damping * gamma + (1.0 - damping) * cp, floor[:, None])
            pred = fwd(mean)
        return dof, nxt

    for it in range(1, t_u + 1):
        dof, nxt = step(it, gamma, noise, True)
        dof = dof.clamp(max=max_dof)
        e = ((yf - pred).abs() ** 2).sum(1)
        cand = torch.maximum((e + noise * dof) / m, floor)
        noise = torch.maximum(damping * noise + (1.0 - damping) * cand, floor)
        gamma = torch.maximum(nxt, floor[:, None])
    step(t_u + 1, gamma, noise, False)

    # polish (solver.py:3961-4511), FP64 on the K atoms
    k = min(polish_size, n)
    score = mean.abs() ** 2 + var
    sup = torch.topk(score, k, dim=1).indices
    g_s = torch.maximum(score.gather(1, sup), floor[:, None]).double()
    d1, d2 = _dirichlet(m1, n1), _dirichlet(m2, n2)
    k1, k2 = sup // n2, sup % n2
    gram = d1[(k1[:, :, None] - k1[:, None, :]) % n1] * d2[(k2[:, :, None] - k2[:, None, :]) % n2]
    cdat = bp.gather(1, sup).to(torch.complex128)
    ye = (yf.abs().double() ** 2).sum(1)
    nu = noise.double()
    fl = floor.double()
    eye = torch.eye(k, dtype=torch.complex128)
    best = None
    for pit in range(evaluations):
        system = torch.diag_embed(1.0 / g_s).to(torch.complex128) + gram / nu[:, None, None]
        system = 0.5 * (system + system.mH)
        ds = system.diagonal(dim1=1, dim2=2).real.clamp_min(tiny).sqrt()
        eq = system / ds[:, :, None] / ds[:, None, :]
        eq = 0.5 * (eq + eq.mH)
        sc = eq.diagonal(dim1=1, dim2=2).real.abs().mean(1).clamp_min(tiny)
        chol = torch.linalg.cholesky(eq + (jitter * eps * sc)[:, None, None] * eye)
        rhs = cdat / nu[:, None]
        mu = torch.cholesky_solve((rhs / ds)[:, :, None], chol)[:, :, 0] / ds
        linv = torch.linalg.solve_triangular(chol, eye.expand(b, k, k), upper=False)
        sig = ((linv.abs() ** 2).sum(1) / ds ** 2).clamp_min(0.0)
        logdet = m * nu.log() + g_s.log().sum(1) + 2.0 * (
            chol.diagonal(dim1=1, dim2=2).real.clamp_min(tiny).log().sum(1) + ds.log().sum(1))
        nle = logdet + (ye / nu - (rhs.conj() * mu).real.sum(1)).clamp_min(0.0)
        cand = dict(mu=mu, sig=sig, g=g_s, nu=nu, nle=nle)
        if best is None:
            best = cand
        else:
            margin = margin_factor * eps * (1.0 + nle.abs() + best["nle"].abs())
            acc = torch.isfinite(nle) & (~torch.isfinite(best["nle"]) | (nle <= best["nle"] - margin))
            best = {key: torch.where(acc[:, None] if best[key].ndim == 2 else acc, cand[key], best[key])
                    for key in best}
        if pit + 1 == evaluations:
            break
        mu_b, sig_b, g_b, nu_b = best["mu"], best["sig"], best["g"], best["nu"]
        energy = (ye - 2.0 * (mu_b.conj() * cdat).real.sum(1)
                  + torch.einsum("bk,bkl,bl->b", mu_b.conj(), gram, mu_b).real).clamp_min(0.0)
        dof = (1.0 - sig_b / g_b.clamp_min(tiny)).clamp(0.0, 1.0).sum(1)
        g_s = torch.maximum(mu_b.abs() ** 2 + sig_b, fl[:, None]).clamp(max=max_precision)
        nu = torch.maximum((energy + nu_b * dof) / m, fl)
    x = torch.zeros(b, n, dtype=torch.complex64)
    x.scatter_(1, sup, best["mu"].to(torch.complex64))
    return dict(x=x.reshape(b, n1, n2), noise=best["nu"].float(), iterations=t_u)
