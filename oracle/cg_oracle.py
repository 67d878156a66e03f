"""Float64 restatement of the covariance-free PCG + Hutchinson path of
``sbl_2d`` with mode="cg" (test oracle).

TEST INFRASTRUCTURE ONLY.  Follows torch_sbl/solver.py:
  probe count rule 1194-1208, Rademacher probes 5979-6000 (the probe stream
  is drawn from torch's CPU generator, as the reference does on CPU),
  main loop 1315-1904 (active-set switch 1316-1341 with 3184-3216, CG
  schedule 1343-1352 with 6086-6118, warm start 1363-1378 with 6121-6138,
  CG posterior 2991-3036, relevance 3156-3181, MacKay 1669-1677, noise
  1737-1768, damping 1483-1487 and 1776-1798, convergence 1800-1904),
  final posterior 2023-2080, PCG 5582-5723, low-rank preconditioner
  5786-5848, reduced active posterior with diffuse-tail loading 3591-3779.
The Toeplitz/BTTB matvec of operators.py:299-371 is evaluated as the
mathematically identical sigma^2 d + A (gamma . A^H d).  An observation
mask (solver.py:986-1024, operators.py:57-75 and 473-496) keeps the
observed samples only: data, probes, residuals and the preconditioner's
factor columns live on those rows, and every "M" of the loop is their
count.

Joint sparsity (joint=True, solver.py:1090-1099, 1157-1159, 1653-1666,
1769-1774, 3039-3153) shares one variance per cell across the snapshots
and updates it by the M-SBL EM rule; with shared noise the posterior is one
PCG solve whose right-hand sides are the B snapshots and the P probes.

Pins: tests/test_oracle_cg.py (PCG against a dense solve, Hutchinson
unbiasedness against the exact diagonal, preconditioner against the dense
Woodbury inverse, the masked operator against the row-restricted dense
matrix, the shared multi-right-hand-side posterior against the dense
solve, end-to-end against the reference's golden runs with and without a
mask and with joint sparsity).
"""

from __future__ import annotations

from dataclasses import dataclass
from math import ceil

import numpy as np

from .fbme_oracle import (F32_EPS, F32_TINY, fourier_columns, prefix_adjoint,
                          prefix_forward, reduced_posterior, gram_closed_form)


@dataclass
class CGConfig:
    """SBLConfig defaults relevant to mode='cg' (solver.py:34-176)."""

    max_iter: int = 200
    min_iter: int = 3
    tol: float = 1e-4
    hyper_shape: float = 1e-8
    hyper_rate: float = 1e-8
    noise_shape: float = 1e-8
    noise_rate: float = 1e-8
    initial_noise_fraction: float = 1e-2
    fixed_noise_variance: float | None = None
    probe_type: str = "rademacher"
    adaptive_probes: bool = False
    initial_hutchinson_probes: int = 16
    spatial_coupling: float = 0.0
    spatial_boundary: str = "zero"
    initialization_blend: float = 0.1
    noise_update: str = "em"
    damping: float = 0.15
    min_variance: float = 1e-12
    max_precision: float = 1e12
    cg_tolerance: float = 1e-5
    cg_initial_tolerance: float = 1e-2
    cg_max_iter: int = 150
    cg_check_interval: int = 8
    cg_fixed_iterations: int = 0
    cg_preconditioner: str = "low_rank"
    cg_preconditioner_rank: int = 32
    cg_damping: float = 0.9
    hutchinson_probes: int = 32
    max_hutchinson_probes: int = 128
    auto_probe_scaling: bool = True
    adaptive_cg: bool = True
    cg_warm_start: bool = True
    cg_min_iter: int = 20
    convergence_patience: int = 2
    convergence_check_interval: int = 2
    active_set: bool = True
    active_set_start: int = 30
    active_set_check_interval: int = 2
    active_set_min_size: int = 32
    active_set_max_size: int = 256
    active_set_tail_fraction: float = 0.05
    seed: int = 0


def probe_count(n: int, m: int, cfg: CGConfig) -> int:
    """solver.py:1194-1208"""
    if not cfg.auto_probe_scaling:
        return cfg.hutchinson_probes
    under = n / m
    rec = 8.0 * under ** 0.5
    if under > 16.0:
        rec = max(rec, 2.0 * under)
    rec = int(rec + 0.999999)
    return min(cfg.max_hutchinson_probes, max(cfg.hutchinson_probes, rec))


def rademacher_probes(batch: int, m: int, count: int, seed: int) -> np.ndarray:
    """Rademacher signs from torch's CPU generator (solver.py:5979-6000)."""
    import torch
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    signs = 2 * torch.randint(0, 2, (batch, m, count), generator=g) - 1
    return signs.numpy().astype(np.float64)


def hadamard_probes(batch: int, shape, count: int, seed: int) -> np.ndarray:
    """Randomised hierarchical Hadamard probes (solver.py:6002-6083): one
    Rademacher sign per sample and block of columns, times columns of the
    Sylvester Hadamard matrix of the smallest power-of-two embedding of
    `shape`, its rows in Morton (bit-interleaved) order of the sample
    coordinates and its columns in Gray-code order within blocks of
    min(8, 2^floor(log2 count)) (one block when count fills the embedding).
    The Hadamard matrix is built by the Sylvester doubling H -> [[H, H],
    [H, -H]], and the Morton codes by explicit loops."""
    import torch
    shape = tuple(int(v) for v in shape)
    m = int(np.prod(shape))
    bits = [max(0, (length - 1).bit_length()) for length in shape]
    embed = 1 << sum(bits)
    if count > embed:
        raise ValueError("Hadamard probe count cannot exceed its embedding.")
    block = count if count == embed else min(8, 1 << max(0, count.bit_length() - 1))
    nblocks = (count + block - 1) // block
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    signs = (2 * torch.randint(0, 2, (batch, m, nblocks), generator=g) - 1).numpy().astype(np.float64)
    codes = np.zeros(m, np.int64)
    for flat in range(m):
        coord, rem = [], flat
        for length in reversed(shape):
            coord.append(rem % length)
            rem //= length
        coord = coord[::-1]
        code, out = 0, 0
        for ib in range(max(bits, default=0)):
            for d, nb in enumerate(bits):
                if ib < nb:
                    code |= ((coord[d] >> ib) & 1) << out
                    out += 1
        codes[flat] = code
    cols = [(j % block) ^ ((j % block) >> 1) for j in range(count)]
    if embed <= 4096:
        h = np.ones((1, 1))
        while h.shape[0] < embed:
            h = np.block([[h, h], [h, -h]])
        had = h[np.ix_(codes, cols)]
    else:   # H[i, j] = (-1)^popcount(i & j), the entries of the same matrix
        had = 1.0 - 2.0 * (np.bitwise_count(codes[:, None] & np.array(cols)[None, :]) & 1)
    return signs[:, :, [j // block for j in range(count)]] * had[None]


def probe_schedule(it: int, cfg: CGConfig, maximum: int) -> int:
    """Probes used at EM iteration `it` (the probe half of _cg_schedule,
    solver.py:6086-6118): with adaptive_probes, initial_hutchinson_probes
    for iterations 1-5, at least min(24, maximum) for 6-15, then all."""
    if not cfg.adaptive_cg or not cfg.adaptive_probes:
        return maximum
    initial = min(cfg.initial_hutchinson_probes, maximum)
    if it <= 5:
        return initial
    if it <= 15:
        return min(maximum, max(initial, min(24, maximum)))
    return maximum


def neighbor_sum(values, grid, boundary):
    """Sum of the four axial neighbours on the grid (_neighbor_sum_2d,
    solver.py:2396-2425), by explicit loops: 'circular' wraps, 'zero' pads."""
    n1, n2 = grid
    f = values.reshape(values.shape[0], n1, n2)
    out = np.zeros_like(f)
    for i in range(n1):
        for j in range(n2):
            acc = np.zeros(f.shape[0])
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                a, c = i + di, j + dj
                if boundary == "circular":
                    acc = acc + f[:, a % n1, c % n2]
                elif 0 <= a < n1 and 0 <= c < n2:
                    acc = acc + f[:, a, c]
            out[:, i, j] = acc
    return out.reshape(values.shape)


def spatial_variance(pattern, grid, cfg):
    """gamma = 1 / clamp(pattern + beta N(pattern)) (solver.py:2662-2672)."""
    eta = pattern + cfg.spatial_coupling * neighbor_sum(pattern, grid, cfg.spatial_boundary)
    return 1.0 / np.clip(eta, 1.0 / cfg.max_precision, cfg.max_precision)


def resize_warm_start(sol, lanes):
    """_resize_warm_start (solver.py:6121-6138): keep the shared leading
    lanes, zeros for new ones."""
    if sol is None or sol.shape[2] == lanes:
        return sol
    out = np.zeros(sol.shape[:2] + (lanes,), sol.dtype)
    k = min(lanes, sol.shape[2])
    out[:, :, :k] = sol[:, :, :k]
    return out


def cg_schedule(it: int, cfg: CGConfig) -> float:
    """Tolerance schedule (solver.py:6086-6118, probes not adaptive)."""
    if not cfg.adaptive_cg:
        return cfg.cg_tolerance
    if it <= 5:
        return max(cfg.cg_initial_tolerance, cfg.cg_tolerance)
    if it <= 15:
        return max(cfg.cg_tolerance,
                   (cfg.cg_initial_tolerance * cfg.cg_tolerance) ** 0.5)
    return cfg.cg_tolerance


class Operator:
    """A, A^H and Q = noise I + A Gamma A^H on [B, M, R] blocks; with an
    observation mask [M1, M2] the sample dimension holds the observed
    samples only (restrict after A, expand before A^H;
    operators.py:57-75, 473-496)."""

    def __init__(self, grid, samples, mask=None):
        self.grid = grid
        self.samples = samples
        self.n = grid[0] * grid[1]
        self.m_full = samples[0] * samples[1]
        if mask is None:
            self.idx = None
            self.m = self.m_full
        else:
            mask = np.asarray(mask, bool)
            if mask.shape != tuple(samples):
                raise ValueError("observation_mask must match the sample shape")
            self.idx = np.flatnonzero(mask.reshape(-1))
            if self.idx.size == 0:
                raise ValueError("observation_mask must retain at least one sample")
            self.m = int(self.idx.size)

    def restrict(self, full):   # [B, M_full, ...] -> [B, M, ...]
        return full if self.idx is None else full[:, self.idx]

    def expand(self, part):     # [B, M, ...] -> [B, M_full, ...]
        if self.idx is None:
            return part
        out = np.zeros((part.shape[0], self.m_full) + part.shape[2:], part.dtype)
        out[:, self.idx] = part
        return out

    def columns(self, indices):   # A[:, S] on the observed rows -> [B, M, K]
        return self.restrict(fourier_columns(indices, self.grid, self.samples))

    def fwd(self, x):      # [B, N, R] -> [B, M, R]
        b, _, r = x.shape
        flat = np.moveaxis(x, 2, 1).reshape(b * r, self.n)
        out = prefix_forward(flat, self.grid, self.samples)
        return self.restrict(np.moveaxis(out.reshape(b, r, self.m_full), 1, 2))

    def adj(self, v):      # [B, M, R] -> [B, N, R]
        v = self.expand(v)
        b, _, r = v.shape
        flat = np.moveaxis(v, 2, 1).reshape(b * r, self.m_full)
        out = prefix_adjoint(flat, self.grid, self.samples)
        return np.moveaxis(out.reshape(b, r, self.n), 1, 2)

    def q(self, d, gamma, noise):
        return noise[:, None, None] * d + self.fwd(gamma[:, :, None] * self.adj(d))


def low_rank_preconditioner(gamma, noise, op: Operator, rank, eps, tiny):
    """Isotropic background plus the strongest atoms (solver.py:5786-5848);
    rank 0 is the scalar background noise + sum gamma (solver.py:5800-5810)."""
    rank = min(rank, op.m, op.n)
    if rank == 0:
        b = gamma.shape[0]
        return noise + gamma.sum(1), np.zeros((b, op.m, 0), np.complex128), np.zeros((b, 0, 0), np.complex128)
    order = np.argsort(-gamma, axis=1, kind="stable")[:, :rank]
    vals = np.take_along_axis(gamma, order, 1)
    cols = op.columns(order)
    factors = cols * np.sqrt(vals)[:, None, :]
    total = gamma.sum(1)
    omitted = np.maximum(total - vals.sum(1), 0.0)
    bg = np.maximum(noise + omitted, eps * total)
    gram = np.einsum("bmk,bml->bkl", np.conj(factors), factors)
    small = np.eye(rank) + gram / bg[:, None, None]
    small = 0.5 * (small + np.conj(np.swapaxes(small, 1, 2)))
    scale = np.maximum(np.mean(np.abs(np.real(np.diagonal(small, axis1=1, axis2=2))), 1), tiny)
    for mult in (10.0, 1e2, 1e3, 1e4, 1e5):
        try:
            chol = np.linalg.cholesky(small + (mult * eps * scale)[:, None, None] * np.eye(rank))
            break
        except np.linalg.LinAlgError:
            chol = None
    linv = np.linalg.inv(chol)
    inv = np.conj(np.swapaxes(linv, 1, 2)) @ linv
    return bg, factors, inv


def circulant_preconditioner(gamma, noise, op: Operator, eps):
    """Strang circulant (BCCB) approximation of Q on the M1 x M2 sample grid
    (operators.py:373-398): the circulant column takes the signed lag
    (k <= M / 2 ? k : k - M) on each axis from fft2(gamma); eigenvalues
    Re fft2(column) + noise, floored at max(eps max|eig|, noise)."""
    b = gamma.shape[0]
    (n1, n2), (m1, m2) = op.grid, op.samples
    lags = np.fft.fft2(gamma.reshape(b, n1, n2), axes=(1, 2))
    k1, k2 = np.arange(m1), np.arange(m2)
    i1 = np.where(k1 <= m1 // 2, k1, n1 - (m1 - k1))
    i2 = np.where(k2 <= m2 // 2, k2, n2 - (m2 - k2))
    eig = np.real(np.fft.fft2(lags[:, i1[:, None], i2[None, :]], axes=(1, 2))) + noise[:, None, None]
    floor = np.maximum(eps * np.abs(eig).reshape(b, -1).max(1), noise)
    return ("circulant", np.maximum(eig, floor[:, None, None]), op)


def grid_diagonal_weights(gamma, noise, order, logscale=None, clamp=4.0):
    """w[b, n] = exp(clamp(o)) / (N (noise + N gamma_T)), gamma_T = gamma
    without the atoms `order` [B, K] (no reference counterpart: the scalar
    background noise + omitted of solver.py:5813-5819 taken per grid cell).
    For M = N samples B^-1 = A diag(w) A^H is exactly Q^-1 of the variances
    gamma_T (A = F, F F^H = N I)."""
    b, n = gamma.shape
    gt = gamma.astype(np.float64).copy()
    gt[np.arange(b)[:, None], order] = 0.0
    w = 1.0 / (n * (noise[:, None] + n * gt))
    if logscale is not None:
        w = w * np.exp(np.clip(np.asarray(logscale, np.float64).reshape(b, n), -clamp, clamp))
    return w


def grid_diagonal_preconditioner(gamma, noise, op: Operator, rank=32, logscale=None, clamp=4.0):
    """(B + F F^H)^-1 with B^-1 = A diag(w) A^H (grid_diagonal_weights) and F
    the `rank` strongest atoms scaled by sqrt(gamma): the low-rank
    preconditioner of solver.py:5786-5848 with its isotropic background
    replaced.  Applied in the Woodbury form B^-1 - U C^-1 U^H, U = B^-1 F,
    C = I + F^H B^-1 F (Hermitian positive definite for any w > 0)."""
    if op.idx is not None:
        raise NotImplementedError("the grid-diagonal background takes unmasked samples")
    order = np.argsort(-gamma, axis=1, kind="stable")[:, :rank]
    vals = np.take_along_axis(gamma, order, 1)
    w = grid_diagonal_weights(gamma, noise, order, logscale, clamp)
    factors = op.columns(order) * np.sqrt(np.maximum(vals, 0.0))[:, None, :]      # [B, M, K]

    def binv(v):                                                                    # [B, M, R]
        return op.fwd(w[:, :, None] * op.adj(v))

    u = binv(factors)
    c = np.eye(rank) + np.einsum("bmk,bml->bkl", np.conj(factors), u)
    ci = np.linalg.inv(0.5 * (c + np.conj(np.swapaxes(c, 1, 2))))
    return ("grid_diagonal", binv, u, ci)


def nystrom_random(batch: int, m: int, count: int, seed: int) -> np.ndarray:
    """The Gaussian test vectors of the Nystrom preconditioner: torch.randn
    real parts, then imaginary parts, [batch, m, count] from torch's CPU
    generator seeded with `seed` (solver.py:5880-5902)."""
    import torch
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    re = torch.randn(batch, m, count, generator=g, dtype=torch.float32).numpy()
    im = torch.randn(batch, m, count, generator=g, dtype=torch.float32).numpy()
    return re.astype(np.float64) + 1j * im.astype(np.float64)


def nystrom_preconditioner(gamma, noise, op: Operator, rank, seed, eps, tiny):
    """Randomised Nystrom approximation of A Gamma A^H (solver.py:5851-5956):
    test matrix = the rank // 2 strongest atoms' columns and rank - rank // 2
    Gaussian vectors, orthonormalised (QR); Y = A Gamma A^H Q, core = Q^H Y
    (symmetrised, jitter-loaded Cholesky L), Y L^-H = U S V^H (SVD),
    eigenvalues S^2; the background is the noise plus the trace the
    approximation misses per sample.  Returns ("nystrom", bg, U, lambda)."""
    b = gamma.shape[0]
    rank = min(rank, op.m)
    total = gamma.sum(1)
    empty = ("nystrom", noise + total, np.zeros((b, op.m, 0), np.complex128), np.zeros((b, 0)))
    if rank == 0:
        return empty
    informed = min(rank, max(1, rank // 2))
    order = np.argsort(-gamma, axis=1, kind="stable")[:, :informed]
    omega = op.columns(order)
    if rank > informed:
        omega = np.concatenate([omega, nystrom_random(b, op.m, rank - informed, seed)], axis=2)
    q = np.linalg.qr(omega)[0]
    y = op.fwd(gamma[:, :, None] * op.adj(q))
    core = np.conj(np.swapaxes(q, 1, 2)) @ y
    core = 0.5 * (core + np.conj(np.swapaxes(core, 1, 2)))
    scale = np.maximum(np.mean(np.abs(np.real(np.diagonal(core, axis1=1, axis2=2))), 1), tiny)
    chol = None
    for mult in (10.0, 1e2, 1e3, 1e4, 1e5):
        try:
            chol = np.linalg.cholesky(core + (mult * eps * scale)[:, None, None] * np.eye(rank))
            break
        except np.linalg.LinAlgError:
            chol = None
    if chol is None:
        return empty
    factor = np.conj(np.swapaxes(np.linalg.solve(chol, np.conj(np.swapaxes(y, 1, 2))), 1, 2))
    u, s, _ = np.linalg.svd(factor, full_matrices=False)
    lam = s ** 2
    resid = np.maximum(total - lam.sum(1) / op.m, 0.0)
    bg = np.maximum(noise + resid, tiny)
    return ("nystrom", bg, u, lam)


def make_preconditioner(kind, gamma, noise, op: Operator, rank, eps, tiny, seed=0):
    """_make_pcg_preconditioner (solver.py:5726-5762) for 'low_rank',
    'nystrom', 'circulant' and 'none'; 'grid_diagonal' is
    grid_diagonal_preconditioner."""
    if kind == "none":
        return ("none",)
    if kind == "nystrom":
        return nystrom_preconditioner(gamma, noise, op, rank, seed, eps, tiny)
    if kind == "grid_diagonal":
        return grid_diagonal_preconditioner(gamma, noise, op, rank)
    if kind == "circulant":
        return circulant_preconditioner(gamma, noise, op, eps)
    return low_rank_preconditioner(gamma, noise, op, rank, eps, tiny)


def apply_preconditioner(v, pre):
    """_apply_pcg_preconditioner (solver.py:5765-5783); the circulant form
    is operators.py:399-416 (expand, transform, divide, invert, restrict)."""
    kind = pre[0] if isinstance(pre[0], str) else "low_rank"
    if kind == "none":
        return v.copy()
    if kind == "nystrom":
        # Frangella-Tropp-Udell flattening (solver.py:5959-5976)
        _, bg, u, lam = pre
        if u.shape[2] == 0:
            return v / bg[:, None, None]
        proj = np.conj(np.swapaxes(u, 1, 2)) @ v
        w = (lam[:, -1:] + bg[:, None]) / (bg[:, None] + lam)
        return v + u @ ((w - 1.0)[:, :, None] * proj)
    if kind == "grid_diagonal":
        _, binv, u, ci = pre
        proj = np.conj(np.swapaxes(u, 1, 2)) @ v
        return binv(v) - u @ (ci @ proj)
    if kind == "circulant":
        _, eig, op = pre
        b, _, r = v.shape
        (m1, m2) = op.samples
        field = op.expand(v).reshape(b, m1, m2, r)
        solved = np.fft.ifft2(np.fft.fft2(field, axes=(1, 2)) / eig[..., None], axes=(1, 2))
        return op.restrict(solved.reshape(b, m1 * m2, r))
    bg, factors, inv = pre
    base = v / bg[:, None, None]
    proj = np.conj(np.swapaxes(factors, 1, 2)) @ v
    corr = inv @ (proj / bg[:, None, None])
    return base - (factors @ corr) / bg[:, None, None]


def pcg(rhs, gamma, noise, op: Operator, *, x0, tol, max_iter, check, rank,
        fixed=0, eps=F32_EPS, tiny=F32_TINY, precond="low_rank", seed=0):
    """Batched complex PCG with lane masks and true-residual restarts
    (solver.py:5582-5723)."""
    if x0 is None:
        x = np.zeros_like(rhs)
        r = rhs.copy()
    else:
        x = x0.copy()
        r = rhs - op.q(x, gamma, noise)
    pre = make_preconditioner(precond, gamma, noise, op, rank, eps, tiny, seed)
    z = apply_preconditioner(r, pre)
    d = z.copy()
    rho = np.real(np.sum(np.conj(r) * z, axis=1))
    init = np.maximum(np.linalg.norm(rhs, axis=1), eps)
    best_x = x.copy() if fixed > 0 else None
    best_rel = np.linalg.norm(r, axis=1) / init if fixed > 0 else None
    active = np.linalg.norm(r, axis=1) > tol * init
    d = np.where(active[:, None, :], d, 0.0)
    limit = fixed if fixed > 0 else max_iter
    it = 0
    for it in range(1, limit + 1):
        qd = op.q(d, gamma, noise)
        curv = np.real(np.sum(np.conj(d) * qd, axis=1))
        valid = (active & np.isfinite(curv) & np.isfinite(rho)
                 & (curv > tiny) & (rho > tiny))
        alpha = np.where(valid, rho / np.maximum(curv, tiny), 0.0)
        x = x + alpha[:, None, :] * d
        r = r - alpha[:, None, :] * qd
        active = valid
        if it % check == 0 or it == limit:
            r = rhs - op.q(x, gamma, noise)
            rel = np.linalg.norm(r, axis=1) / init
            if fixed > 0:
                better = rel < best_rel
                best_x = np.where(better[:, None, :], x, best_x)
                best_rel = np.minimum(best_rel, rel)
            ra = np.isfinite(rel) & (rel > tol)
            active = (active & ra) if fixed > 0 else ra
            z = apply_preconditioner(r, pre)
            rz = np.real(np.sum(np.conj(r) * z, axis=1))
            active = active & np.isfinite(rz) & (rz > tiny)
            d = np.where(active[:, None, :], z, 0.0)
            rho = np.where(active, rz, rho)
            if fixed == 0 and not active.any():
                break
            continue
        z = apply_preconditioner(r, pre)
        nrho = np.real(np.sum(np.conj(r) * z, axis=1))
        vu = active & np.isfinite(nrho) & (nrho > tiny)
        beta = np.where(vu, nrho / np.maximum(rho, tiny), 0.0)
        d = np.where(vu[:, None, :], z + beta[:, None, :] * d, 0.0)
        rho = np.where(vu, nrho, rho)
        active = vu
    return (best_x if best_x is not None else x), it


def project_relevance(gamma, dd, m, eps):
    rel = np.clip(gamma * dd, 0.0, 1.0)
    raw = rel.sum(1)
    scale = np.minimum(1.0, m * (1.0 - 10.0 * eps) / np.maximum(raw, eps))
    return rel * scale[:, None]


def select_active(gamma, cfg: CGConfig):
    """solver.py:3184-3216 (batch of one)."""
    if gamma.shape[0] != 1 or not np.all(np.isfinite(gamma)):
        return None
    size_max = min(cfg.active_set_max_size, gamma.shape[1])
    order = np.argsort(-gamma[0], kind="stable")[:size_max]
    vals = gamma[0, order]
    total = gamma[0].sum()
    target = (1.0 - cfg.active_set_tail_fraction) * total
    if vals.sum() < target:
        return None
    cum = np.cumsum(vals)
    required = min(int(np.searchsorted(cum, target)), size_max - 1) + 1
    size = min(size_max, max(cfg.active_set_min_size, required))
    return order[None, :size]


def active_posterior(y, gamma, noise, sup, op: Operator, eps, tiny):
    """Reduced posterior with diffuse-tail loading nu = noise + omitted
    variance, CPU Cholesky loading (solver.py:3591-3779, 3897-3924)."""
    b = y.shape[0]
    rows = np.arange(b)[:, None]
    gs = gamma[rows, sup]
    omitted = np.maximum(gamma.sum(1) - gs.sum(1), 0.0)
    nu = noise + omitted
    if op.idx is None:
        gram = gram_closed_form(sup, op.grid, op.samples)
    else:
        cols = op.columns(sup)
        gram = np.einsum("bmk,bml->bkl", np.conj(cols), cols)
    cdat = op.adj(y[:, :, None])[:, :, 0][rows, sup]
    mu, sig, dd, _ = reduced_posterior(gram, cdat, np.sum(np.abs(y) ** 2, 1),
                                       gs, nu, op.m, eps=eps, tiny=tiny,
                                       jitter_multiplier=1.0)
    mean = np.zeros_like(gamma, dtype=np.complex128)
    mean[rows, sup] = mu
    var = gamma.copy()
    var[rows, sup] = sig
    ddiag = np.zeros_like(gamma)
    ddiag[rows, sup] = dd
    return mean, var, ddiag


def _posterior_shared(y, gamma, noise, probes, lprobes, op, sol, cfg, tol, eps, tiny):
    """One system for all snapshots (shared gamma and noise): the B data
    vectors and the P probes are the right-hand sides of one PCG solve
    (_shared_mmv_posterior, solver.py:3039-3153)."""
    b = y.shape[0]
    rhs = np.concatenate([y.T[None], probes], axis=2)           # [1, M, B + P]
    sol = resize_warm_start(sol, rhs.shape[2])
    sol, nit = pcg(rhs, gamma[:1], noise[:1], op, x0=sol if cfg.cg_warm_start else None,
                   tol=tol, max_iter=cfg.cg_max_iter, check=cfg.cg_check_interval,
                   rank=cfg.cg_preconditioner_rank, fixed=cfg.cg_fixed_iterations, eps=eps, tiny=tiny,
                   precond=cfg.cg_preconditioner, seed=cfg.seed)
    lifted = op.adj(sol)                                        # [1, N, B + P]
    dd1 = np.maximum(np.mean(np.real(lifted[:, :, b:] * np.conj(lprobes)), 2), 0.0)   # [1, N]
    mean = gamma[:1] * lifted[0, :, :b].T                       # [B, N]
    var1 = np.maximum(gamma[:1] - gamma[:1] ** 2 * dd1, 0.0)
    return mean, np.repeat(var1, b, 0), np.repeat(dd1, b, 0), sol, nit


def sbl_cg(data, grid, cfg: CGConfig | None = None, *, eps=F32_EPS,
           tiny=F32_TINY, trace=None, mask=None, joint=False, joint_noise=True) -> dict:
    """Default ``sbl_2d`` in mode='cg' on [B, M1, M2] (or [M1, M2]) data,
    optionally on the samples an observation mask [M1, M2] keeps.

    joint: joint sparsity (M-SBL, solver.py:1090-1099, 1157-1159,
    1653-1666, 1769-1774): one variance per grid cell for every snapshot,
    initialised to the snapshots' mean, floored at the largest floor and
    updated by the EM rule on the snapshots' mean second moment; with
    joint_noise the noise variance is shared too and the posterior is one
    multi-right-hand-side solve."""
    cfg = cfg or CGConfig()
    data = np.asarray(data)
    single = data.ndim == 2
    if single:
        data = data[None]
    b = data.shape[0]
    samples = tuple(data.shape[-2:])
    grid = tuple(int(v) for v in grid)
    op = Operator(grid, samples, mask)
    n, m = op.n, op.m
    y = op.restrict(data.reshape(b, op.m_full).astype(np.complex128))
    power = np.maximum(np.mean(np.abs(y) ** 2, 1), eps)
    floor = np.maximum(power * cfg.min_variance, tiny)
    bp = op.adj(y[:, :, None])[:, :, 0]
    energy = np.abs(bp) ** 2
    structured = energy / np.maximum(energy.sum(1, keepdims=True), eps) * power[:, None]
    gamma = (cfg.initialization_blend * structured
             + (1 - cfg.initialization_blend) * power[:, None] / n)
    shared = joint and joint_noise
    if joint:
        if b < 2:
            raise ValueError("joint_sparsity=True requires at least two leading batch items.")
        gamma = np.repeat(gamma.mean(0, keepdims=True), b, 0)
        prior_floor = np.full(b, floor.max())
    else:
        prior_floor = floor
    gamma = np.maximum(gamma, prior_floor[:, None])
    pattern = None
    if cfg.spatial_coupling > 0.0:
        # scale-matched base precision of the coupled prior (solver.py:1121-1150, 2674-2687)
        degree = neighbor_sum(np.ones((1, n)), grid, cfg.spatial_boundary)
        pattern = (1.0 / gamma) / (1.0 + cfg.spatial_coupling * degree)
        gamma = np.maximum(spatial_variance(pattern, grid, cfg), prior_floor[:, None])
    noise = np.maximum(cfg.initial_noise_fraction * power, floor)
    if cfg.fixed_noise_variance is not None:
        noise = np.full(b, cfg.fixed_noise_variance)      # solver.py:1151-1156
    if shared:
        noise = np.full(b, noise.mean())
    pcount = probe_count(n, m, cfg)
    if cfg.probe_type == "hadamard":
        # probe_shape: the sample grid, or the observed samples flat under a
        # mask (solver.py:1209-1220)
        pshape = op.samples if op.idx is None else (m,)
        probes = hadamard_probes(1 if shared else b, pshape, pcount, cfg.seed).astype(np.complex128)
    else:
        probes = rademacher_probes(1 if shared else b, m, pcount, cfg.seed).astype(np.complex128)
    lprobes = op.adj(probes)
    prev = np.zeros((b, n), np.complex128)
    keep = max(cfg.damping, cfg.cg_damping)
    sup = None
    sol = None
    streak = 0
    converged = False
    min_iters = max(cfg.min_iter, cfg.cg_min_iter)
    tol_floor = 3.0 * eps ** 0.5
    it = 0
    cg_iters = []
    for it in range(1, cfg.max_iter + 1):
        if (cfg.active_set and b == 1 and not joint and sup is None and pattern is None
                and it >= cfg.active_set_start
                and (it - cfg.active_set_start) % cfg.active_set_check_interval == 0):
            sup = select_active(gamma, cfg)
            if sup is not None:
                sol = None
        tol = max(cg_schedule(it, cfg), tol_floor)
        used = probe_schedule(it, cfg, pcount)
        if sup is not None:
            mean, var, dd = active_posterior(y, gamma, noise, sup, op, eps, tiny)
            cg_iters.append(0)
            sol = None
        elif shared:
            mean, var, dd, sol, nit = _posterior_shared(y, gamma, noise, probes[:, :, :used],
                                                        lprobes[:, :, :used], op, sol, cfg, tol, eps, tiny)
            cg_iters.append(nit)
        else:
            rhs = np.concatenate([y[:, :, None], probes[:, :, :used]], axis=2)
            sol = resize_warm_start(sol, rhs.shape[2])
            sol, nit = pcg(rhs, gamma, noise, op, x0=sol if cfg.cg_warm_start else None,
                           tol=tol, max_iter=cfg.cg_max_iter,
                           check=cfg.cg_check_interval,
                           rank=cfg.cg_preconditioner_rank, precond=cfg.cg_preconditioner,
                           fixed=cfg.cg_fixed_iterations, eps=eps, tiny=tiny, seed=cfg.seed)
            cg_iters.append(nit)
            lifted = op.adj(sol)
            dd = np.maximum(np.mean(np.real(lifted[:, :, 1:] * np.conj(lprobes[:, :, :used])), 2), 0.0)
            mean = gamma * lifted[:, :, 0]
            var = np.maximum(gamma - gamma ** 2 * dd, 0.0)
        rel = project_relevance(gamma, dd, m, eps)
        dof = rel.sum(1)
        if pattern is not None:
            # pattern-coupled update (solver.py:1489-1590, 2690-2720)
            second = np.abs(mean) ** 2 + var
            coupled = second + cfg.spatial_coupling * neighbor_sum(second, grid, cfg.spatial_boundary)
            base = np.clip((1.0 + cfg.hyper_shape) / (coupled + cfg.hyper_rate),
                           1.0 / cfg.max_precision, cfg.max_precision)
            pattern = keep * pattern + (1 - keep) * base
            new_gamma = spatial_variance(pattern, grid, cfg)
        elif joint:
            second = np.mean(np.abs(mean) ** 2 + var, 0, keepdims=True)
            new_gamma = np.repeat(np.clip((second + cfg.hyper_rate) / (1.0 + cfg.hyper_shape),
                                          prior_floor.max(), cfg.max_precision), b, 0)
        else:
            prec = np.clip((rel + cfg.hyper_shape) / (np.abs(mean) ** 2 + cfg.hyper_rate),
                           1.0 / cfg.max_precision, cfg.max_precision)
            new_gamma = 1.0 / prec
        if sup is not None:
            mask = np.zeros_like(gamma, bool)
            mask[np.arange(b)[:, None], sup] = True
            new_gamma = np.where(mask, new_gamma, gamma)
        resid = y - op.fwd(mean[:, :, None])[:, :, 0]
        e = np.sum(np.abs(resid) ** 2, 1)
        if cfg.noise_update == "em":
            new_noise = (e + noise * dof + cfg.noise_rate) / (m + cfg.noise_shape)
        else:
            new_noise = (e + cfg.noise_rate) / np.maximum(m - dof + cfg.noise_shape, eps)
        if cfg.fixed_noise_variance is not None:
            new_noise = np.full(b, cfg.fixed_noise_variance)   # solver.py:1747-1750
        new_noise = np.maximum(new_noise, floor)
        if shared:
            new_noise = np.full(b, max(new_noise.mean(), floor.max()))
        if pattern is not None:
            gamma = np.maximum(new_gamma, prior_floor[:, None])   # damped in the pattern
        else:
            gamma = np.maximum(keep * gamma + (1 - keep) * new_gamma, prior_floor[:, None])
        if cfg.fixed_noise_variance is not None:
            noise = new_noise                                  # undamped (solver.py:1788-1798)
        else:
            noise = keep * noise + (1 - keep) * new_noise
        change = np.max(np.linalg.norm(mean - prev, axis=1)
                        / np.maximum(np.linalg.norm(mean, axis=1), eps))
        prev = mean
        if trace is not None:
            trace.append(dict(change=change, noise=noise.copy(), active=0 if sup is None else sup.shape[1]))
        check = (cfg.tol > 0 and it >= min_iters
                 and ((it - min_iters) % cfg.convergence_check_interval == 0
                      or it == cfg.max_iter))
        if check:
            streak = streak + 1 if change <= cfg.tol else 0
            if streak >= cfg.convergence_patience:
                converged = True
                break
    final_tol = max(cfg.cg_tolerance, tol_floor)
    if sup is not None:
        mean, var, dd = active_posterior(y, gamma, noise, sup, op, eps, tiny)
    elif shared:
        mean, var, dd, sol, _ = _posterior_shared(y, gamma, noise, probes, lprobes, op, sol, cfg,
                                                  final_tol, eps, tiny)
    else:
        rhs = np.concatenate([y[:, :, None], probes], axis=2)
        sol, _ = pcg(rhs, gamma, noise, op, x0=sol if cfg.cg_warm_start else None,
                     tol=final_tol, max_iter=cfg.cg_max_iter,
                     check=cfg.cg_check_interval, rank=cfg.cg_preconditioner_rank, precond=cfg.cg_preconditioner,
                     fixed=cfg.cg_fixed_iterations, eps=eps, tiny=tiny)
        lifted = op.adj(sol)
        dd = np.maximum(np.mean(np.real(lifted[:, :, 1:] * np.conj(lprobes)), 2), 0.0)
        mean = gamma * lifted[:, :, 0]
        var = np.maximum(gamma - gamma ** 2 * dd, 0.0)
    dof = project_relevance(gamma, dd, m, eps).sum(1)
    out = dict(x=mean, posterior_variance=var, precision=1.0 / gamma,
               noise_variance=noise, effective_dof=dof, iterations=it,
               converged=converged, active_set_size=0 if sup is None else sup.shape[1],
               cg_iterations=cg_iters, probes=pcount)
    if single:
        out = {k: (v[0] if isinstance(v, np.ndarray) and v.shape[:1] == (1,) else v)
               for k, v in out.items()}
    return out
