"""Float64 restatement of the off-grid refinement of an on-grid SBL estimate
(test oracle).

TEST INFRASTRUCTURE ONLY.  Follows torch_sbl/offgrid.py:
  continuous Fourier atoms 29-67, driver 70-226 (peak selection, optional
  BIC model order, one refinement per batch item), peak selection with
  periodic non-maximum suppression 229-295, complex-data extended BIC
  318-361, projected Gauss--Newton with Levenberg damping, trust bounds and
  monotone backtracking 364-476, ridge-regularised amplitude fit 479-499,
  projected position Jacobian 502-535, position standard deviation from the
  Fisher matrix 538-564.

Everything is float64.  The reference evaluates in the data's dtype
(float32 for complex64 data); the eps / tiny constants it takes from that
dtype are used here as given by `eps` / `tiny`.

Pins: tests/test_oracle_offgrid.py (atoms against the prefix DFT at integer
positions, Jacobian against finite differences, the refinement recovering
planted off-grid frequencies, and the reference's golden runs).
"""

from __future__ import annotations

from math import log

import numpy as np

F32_EPS = float(np.finfo(np.float32).eps)
F32_TINY = float(np.finfo(np.float32).tiny)


def sample_coordinates(samples):
    """[M, D] sample coordinates in row-major order (offgrid.py:49-58)."""
    axes = [np.arange(m, dtype=np.float64) for m in samples]
    return np.stack([g.reshape(-1) for g in np.meshgrid(*axes, indexing="ij")], axis=1)


def continuous_dictionary(positions, grid, samples):
    """[M, K] atoms exp(-2 pi i sum_d m_d p_d / N_d) (offgrid.py:29-67)."""
    coords = sample_coordinates(samples) / np.asarray(grid, np.float64)[None, :]
    return np.exp(-2j * np.pi * (coords @ np.asarray(positions, np.float64).T))


def select_peaks(magnitude, grid, component_count, max_components, relative_threshold,
                 minimum_separation):
    """Greedy periodic non-maximum suppression over the strongest cells
    (offgrid.py:229-295).  Ties in magnitude go to the lower flat index."""
    n = magnitude.size
    count = min(n, max(64, 16 * max_components))
    order = np.lexsort((np.arange(n), -magnitude))[:count]
    values = magnitude[order]
    cutoff = relative_threshold * values[0]
    g = np.asarray(grid, np.float64)
    selected, chosen = [], set()
    target = component_count if component_count is not None else max_components
    for off in range(count):
        if component_count is None and selected and values[off] < cutoff:
            break
        idx = int(order[off])
        coord = np.array(divmod(idx, grid[1]), np.float64) if len(grid) == 2 else np.array([idx], np.float64)
        if selected:
            diff = np.abs(np.stack(selected) - coord)
            dist = np.sqrt(np.sum(np.minimum(diff, g[None, :] - diff) ** 2, axis=1))
            if np.any(dist < minimum_separation):
                continue
        selected.append(coord)
        chosen.add(idx)
        if len(selected) >= target:
            break
    if component_count is not None and len(selected) < component_count:
        for idx in order:
            idx = int(idx)
            if idx in chosen:
                continue
            selected.append(np.array(divmod(idx, grid[1]), np.float64))
            chosen.add(idx)
            if len(selected) == component_count:
                break
    if not selected:
        selected.append(np.array(divmod(int(order[0]), grid[1]), np.float64))
    return np.stack(selected)


def bic_order(y, positions, grid, samples, eps=F32_EPS):
    """Number of leading peaks minimising the extended complex BIC
    (offgrid.py:318-361); orthogonalisation by QR of the atoms."""
    d = continuous_dictionary(positions, grid, samples)
    q, _ = np.linalg.qr(d)
    proj = np.abs(q.conj().T @ y) ** 2
    resid = np.maximum(np.sum(np.abs(y) ** 2) - np.cumsum(proj), eps)
    k = np.arange(1, positions.shape[0] + 1, dtype=np.float64)
    m = float(y.size)
    bic = (m * np.log(resid / m) + (2 + len(grid)) * k * log(m)
           + 2.0 * k * log(float(np.prod(grid))))
    return int(np.argmin(bic)) + 1


def fit_amplitudes(d, y, ridge, eps=F32_EPS, tiny=F32_TINY):
    """Ridge-loaded least squares (offgrid.py:479-499): returns the
    amplitudes and the Cholesky factor of the loaded Gram."""
    gram = d.conj().T @ d
    mean_diag = max(float(np.mean(np.real(np.diag(gram)))), tiny)
    gram = gram + max(ridge, eps) * mean_diag * np.eye(gram.shape[0])
    gram = 0.5 * (gram + gram.conj().T)
    chol = np.linalg.cholesky(gram)
    rhs = d.conj().T @ y
    amp = np.linalg.solve(chol.conj().T, np.linalg.solve(chol, rhs))
    return amp, chol


def projected_jacobian(d, amp, chol, grid, samples):
    """dr/dp of the variable-projection residual, first-order (Kaufman)
    form (offgrid.py:502-535): [M, K * D], column (k, dim)."""
    coords = sample_coordinates(samples) / np.asarray(grid, np.float64)[None, :]
    factor = -2j * np.pi * coords                                   # [M, D]
    jac = (d[:, :, None] * factor[:, None, :] * amp[None, :, None]).reshape(d.shape[0], -1)
    coef = np.linalg.solve(chol.conj().T, np.linalg.solve(chol, d.conj().T @ jac))
    return jac - d @ coef


def refine_single(y, positions, grid, samples, *, max_iter, tolerance, max_step, damping,
                  ridge, eps=F32_EPS, tiny=F32_TINY):
    """Projected Gauss--Newton with Levenberg damping, per-coordinate trust
    bound and monotone backtracking (offgrid.py:364-476)."""
    pos = np.array(positions, np.float64)
    g = np.asarray(grid, np.float64)
    lam = damping
    d = continuous_dictionary(pos, grid, samples)
    amp, chol = fit_amplitudes(d, y, ridge, eps, tiny)
    resid = y - d @ amp
    initial_norm = float(np.linalg.norm(resid))
    obj = float(np.sum(np.abs(resid) ** 2))
    converged = False
    it = 0
    for it in range(1, max_iter + 1):
        jac = projected_jacobian(d, amp, chol, grid, samples)
        normal = np.real(jac.conj().T @ jac)
        rhs = np.real(jac.conj().T @ resid)
        diag = np.maximum(np.diag(normal), eps)
        step = np.linalg.solve(normal + lam * np.diag(diag), rhs)
        step = np.clip(step.reshape(pos.shape), -max_step, max_step)
        if not np.all(np.isfinite(step)):
            break
        accepted = False
        taken = step
        previous = obj
        for scale in (1.0, 0.5, 0.25, 0.125, 0.0625):
            trial = np.mod(pos + scale * step, g)
            td = continuous_dictionary(trial, grid, samples)
            tamp, tchol = fit_amplitudes(td, y, ridge, eps, tiny)
            tres = y - td @ tamp
            tobj = float(np.sum(np.abs(tres) ** 2))
            if tobj < obj:
                pos, d, amp, chol, resid, obj = trial, td, tamp, tchol, tres, tobj
                taken = scale * step
                accepted = True
                lam = max(lam * 0.5, 1e-8)
                break
        if not accepted:
            lam = min(lam * 10.0, 1e8)
            if lam >= 1e8:
                converged = True
                break
            continue
        rel = (previous - obj) / max(previous, eps)
        if float(np.max(np.abs(taken))) <= tolerance or rel <= tolerance ** 2:
            converged = True
            break
    jac = projected_jacobian(d, amp, chol, grid, samples)
    fisher = np.real(jac.conj().T @ jac)
    dof = max(d.shape[0] - d.shape[1], 1)
    noise = float(np.sum(np.abs(resid) ** 2)) / dof
    cov = 0.5 * noise * np.linalg.pinv(fisher, hermitian=True)
    std = np.sqrt(np.maximum(np.diag(cov), 0.0)).reshape(amp.shape[0], len(grid))
    return dict(positions=pos, amplitudes=amp, position_std=std, initial_residual_norm=initial_norm,
                residual_norm=float(np.linalg.norm(resid)), iterations=it, converged=converged)


def refine_offgrid(data, coefficients, grid, *, component_count=None, max_components=64,
                   relative_threshold=0.1, minimum_separation=1.5, max_iter=15, tolerance=1e-4,
                   max_step=0.35, damping=1e-3, ridge=1e-7, eps=F32_EPS, tiny=F32_TINY):
    """refine_offgrid_fourier on [B, M1, M2] data and [B, N1, N2]
    coefficients (offgrid.py:70-226); returns a list of per-item dicts."""
    grid = tuple(int(v) for v in grid)
    data = np.asarray(data)
    coef = np.asarray(coefficients)
    if data.ndim == len(grid):
        data, coef = data[None], coef[None]
    samples = tuple(int(v) for v in data.shape[-len(grid):])
    b = data.shape[0]
    y = data.reshape(b, -1).astype(np.complex128)
    x = coef.reshape(b, -1)
    maximum = min(max_components, y.shape[1] - 1, x.shape[1])
    if component_count is not None:
        maximum = min(maximum, component_count)
    out = []
    for i in range(b):
        init = select_peaks(np.abs(x[i]).astype(np.float64), grid, component_count, maximum,
                            relative_threshold, minimum_separation)
        if component_count is None and init.shape[0] > 1:
            init = init[:bic_order(y[i], init, grid, samples, eps)]
        res = refine_single(y[i], init, grid, samples, max_iter=max_iter, tolerance=tolerance,
                            max_step=max_step, damping=damping, ridge=ridge, eps=eps, tiny=tiny)
        res["initial_positions"] = init
        out.append(res)
    return out
