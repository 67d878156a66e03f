"""Exact dense EM-SBL in float64 -- the reference's oracle solver
(literature_reproduction/baselines.py:254-278 running torch_sbl.sbl_2d with
mode="dense"; posterior solver.py:2962-2990, covariance operators.py:282-297,
data diagonal operators.py:418-442, MacKay update solver.py:1669-1677,
evidence noise update solver.py:1759-1767).

TEST INFRASTRUCTURE ONLY.  Pinned by tests/test_oracle_dense.py against the
reference's golden vectors and against a direct textbook EM step.
"""

from __future__ import annotations

import numpy as np

from .fbme_oracle import dense_prefix_matrix, prefix_forward

F64_EPS = float(np.finfo(np.float64).eps)
F64_TINY = float(np.finfo(np.float64).tiny)


def dense_posterior(a, y, gamma, noise, eps=F64_EPS):
    """mu = Gamma A^H Q^-1 y, v = gamma - gamma^2 diag(A^H Q^-1 A) with
    Q = noise I + A Gamma A^H; CPU Cholesky with the reference's first
    loading eps * mean|diag| (solver.py:2921-2931)."""
    m = a.shape[0]
    q = (a * gamma[None, :]) @ np.conj(a.T) + noise * np.eye(m)
    scale = max(np.mean(np.abs(np.real(np.diag(q)))), F64_TINY)
    for mult in (1.0, 10.0, 100.0, 1000.0):
        try:
            chol = np.linalg.cholesky(q + mult * eps * scale * np.eye(m))
            break
        except np.linalg.LinAlgError:
            chol = None
    if chol is None:
        raise RuntimeError("Data covariance is not numerically positive definite.")
    qinv = np.linalg.inv(chol)
    qinv = np.conj(qinv.T) @ qinv
    qy = qinv @ y
    diag = np.real(np.sum(np.conj(a) * (qinv @ a), axis=0))
    mean = gamma * (np.conj(a.T) @ qy)
    var = np.maximum(gamma - gamma**2 * diag, 0.0)
    return mean, var, diag


def project_relevance(gamma, diag, m, eps):
    rel = np.clip(gamma * diag, 0.0, 1.0)
    raw = rel.sum()
    return rel * min(1.0, m * (1.0 - 10.0 * eps) / max(raw, eps))


def dense_em_sbl(data, grid, iterations=30, *, hyper=1e-6):
    """Classical exact EM-SBL through the observation-domain covariance."""
    y = np.asarray(data).reshape(-1).astype(np.complex128)
    samples = tuple(np.asarray(data).shape[-2:])
    grid = tuple(int(v) for v in grid)
    a = dense_prefix_matrix(grid, samples)
    m, n = a.shape
    eps, tiny = F64_EPS, F64_TINY
    power = max(np.mean(np.abs(y) ** 2), eps)
    floor = max(power * 1e-12, tiny)
    gamma = np.full(n, max(power / n, floor))
    noise = max(1e-2 * power, floor)
    for _ in range(iterations):
        mean, var, diag = dense_posterior(a, y, gamma, noise, eps)
        rel = project_relevance(gamma, diag, m, eps)
        dof = rel.sum()
        prec = np.clip((rel + hyper) / (np.abs(mean) ** 2 + hyper), 1e-12, 1e12)
        new_gamma = 1.0 / prec
        resid = y - prefix_forward(mean[None], grid, samples)[0]
        energy = np.sum(np.abs(resid) ** 2)
        new_noise = max((energy + hyper) / max(m - dof + hyper, eps), floor)
        gamma = np.maximum(new_gamma, floor)
        noise = new_noise
    mean, var, diag = dense_posterior(a, y, gamma, noise, eps)
    dof = project_relevance(gamma, diag, m, eps).sum()
    return dict(x=mean.reshape(grid), posterior_variance=var.reshape(grid),
                precision=(1.0 / gamma).reshape(grid), noise_variance=noise,
                effective_dof=dof, iterations=iterations)
