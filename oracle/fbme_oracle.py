"""Float64 restatement of the FBME-Fast / FBME-Stable plans (test oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product path.

Every quantity is computed in float64/complex128.  The *constants* of the
algorithm (eps, tiny, the dof cap, the evidence margin, the Cholesky loading)
are those of the reference's arithmetic dtype: complex64 input means float32
eps/tiny exactly as ``operator.real_dtype`` selects them in the reference
(torch_sbl/operators.py:52-55, solver.py:1063-1065).

Pins (see tests/test_oracle_*.py):
  * prefix_forward / prefix_adjoint: brute-force dense DFT submatrix, the
    adjoint identity <Ax, v> = <x, A^H v>, A A^H = N I.
  * fourier_columns / dirichlet: dense DFT columns; Gram = C^H C.
  * reduced posterior + evidence: textbook Gaussian posterior and slogdet of
    the dense covariance nu I + C Gamma C^H (matrix determinant lemma).
  * UAMP step: the reference's Torch branch equations restated
    independently in tests (closed-form elementwise maps on random inputs),
    invariants 0 < v, 0 <= dof <= M.
  * end-to-end Fast/Stable: golden vectors written by
    tests/golden/make_golden.py from the reference's CPU Torch backend,
    with and without an observation mask.

An observation mask [M1, M2] (solver.py:986-1024) keeps the observed rows
of the operator (operators.py:57-75, 473-496): data and messages live on
the observed samples, A gathers them after the transform and A^H scatters
them into zeros before it, Gram matrices are formed from the restricted
columns, and M is the observed count throughout.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from math import ceil, log

import numpy as np
import scipy.fft as _sfft

F32_EPS = float(np.finfo(np.float32).eps)
F32_TINY = float(np.finfo(np.float32).tiny)
F64_EPS = float(np.finfo(np.float64).eps)
F64_TINY = float(np.finfo(np.float64).tiny)


# --------------------------------------------------------------------------
# Prefix DFT operator (torch_sbl/operators.py:237-280, 444-466)
# --------------------------------------------------------------------------

def prefix_forward(x: np.ndarray, grid, samples) -> np.ndarray:
    """A x = fft2(x)[:M1, :M2], flattened.  x: [B, N] -> [B, M]
    (operators.py:237-253)."""
    b = x.shape[0]
    field_ = x.reshape(b, *grid).astype(np.complex128)
    spec = _sfft.fft2(field_, axes=(1, 2), workers=-1)
    return spec[:, : samples[0], : samples[1]].reshape(b, -1)


def prefix_adjoint(v: np.ndarray, grid, samples) -> np.ndarray:
    """A^H v = N * ifft2(zero-pad(v), s=grid), flattened.  [B, M] -> [B, N]
    (operators.py:255-280)."""
    b = v.shape[0]
    field_ = v.reshape(b, *samples).astype(np.complex128)
    n = grid[0] * grid[1]
    out = _sfft.ifft2(field_, s=grid, axes=(1, 2), workers=-1) * n
    return out.reshape(b, -1)


def restrict(v: np.ndarray, mask) -> np.ndarray:
    """[B, M1 M2, ...] -> [B, M_observed, ...] (operators.py:473-482)."""
    return v if mask is None else v[:, np.asarray(mask, bool).reshape(-1)]


def expand(v: np.ndarray, mask) -> np.ndarray:
    """[B, M_observed, ...] -> [B, M1 M2, ...], zeros on the unobserved
    samples (operators.py:484-496)."""
    if mask is None:
        return v
    keep = np.asarray(mask, bool).reshape(-1)
    out = np.zeros((v.shape[0], keep.size) + v.shape[2:], np.result_type(v, np.complex128))
    out[:, keep] = v
    return out


def masked_forward(x, grid, samples, mask=None):
    return restrict(prefix_forward(x, grid, samples), mask)


def masked_adjoint(v, grid, samples, mask=None):
    return prefix_adjoint(expand(v, mask), grid, samples)


def masked_gram(indices, grid, samples, mask=None):
    """C^H C of the selected columns on the observed rows (closed form
    without a mask)."""
    if mask is None:
        return gram_closed_form(indices, grid, samples)
    cols = restrict(fourier_columns(indices, grid, samples), mask)
    return np.einsum("bmk,bml->bkl", np.conj(cols), cols)


def dense_prefix_matrix(grid, samples) -> np.ndarray:
    """Explicit A[m, n] = exp(-2 pi i (m1 n1/N1 + m2 n2/N2)) for tiny sizes."""
    n1, n2 = grid
    m1, m2 = samples
    mm1, mm2 = np.meshgrid(np.arange(m1), np.arange(m2), indexing="ij")
    gg1, gg2 = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    ph = (np.outer(mm1.ravel(), gg1.ravel()) / n1
          + np.outer(mm2.ravel(), gg2.ravel()) / n2)
    return np.exp(-2j * np.pi * ph)


def fourier_columns(indices: np.ndarray, grid, samples,
                    phase_dtype=np.float64) -> np.ndarray:
    """Selected columns A[:, S] -> [B, M, K] (operators.py:444-466).

    ``phase_dtype=np.float32`` reproduces the reference's float32 phase and
    angle (no reduction modulo N) for measuring its own rounding; the default
    is exact.
    """
    n1, n2 = grid
    m1, m2 = samples
    k1 = indices // n2
    k2 = indices % n2
    mm1, mm2 = np.meshgrid(np.arange(m1), np.arange(m2), indexing="ij")
    s1 = (mm1.ravel() / n1).astype(phase_dtype)
    s2 = (mm2.ravel() / n2).astype(phase_dtype)
    ph = (s1[None, :, None] * k1[:, None, :].astype(phase_dtype)
          + s2[None, :, None] * k2[:, None, :].astype(phase_dtype))
    # The reference forms the angle -2 pi phase in the phase dtype before the
    # complex exponential; in float32 that rounding (~1e-4 rad at 256^2)
    # dominates its column error.
    angle = (phase_dtype(-2.0 * np.pi) * ph).astype(np.float64)
    return np.exp(1j * angle)


def dirichlet(m: int, n: int) -> np.ndarray:
    """D[d] = sum_{j<m} exp(2 pi i j d / n), d = 0..n-1 (the 1-D Gram kernel
    of a prefix DFT: (C^H C)_{kl} = D1[(k1-l1) mod N1] * D2[(k2-l2) mod N2])."""
    d = np.arange(n)
    j = np.arange(m)
    return np.exp(2j * np.pi * np.outer(d, j) / n).sum(axis=1)


def gram_closed_form(indices: np.ndarray, grid, samples) -> np.ndarray:
    d1 = dirichlet(samples[0], grid[0])
    d2 = dirichlet(samples[1], grid[1])
    k1 = indices // grid[1]
    k2 = indices % grid[1]
    g1 = d1[(k1[:, :, None] - k1[:, None, :]) % grid[0]]
    g2 = d2[(k2[:, :, None] - k2[:, None, :]) % grid[1]]
    return g1 * g2


# --------------------------------------------------------------------------
# Configuration (subset of SBLConfig used by the locked plans)
# --------------------------------------------------------------------------

@dataclass
class PlanConfig:
    """Numerics of SBLConfig.paper_hybrid_uamp (solver.py:198-280) with the
    Fast (profiles.py:63-70) or Stable (profiles.py:110-115) overrides."""

    uamp_damping: float = 0.05
    damping: float = 0.15
    min_variance: float = 1e-12
    max_precision: float = 1e12
    initialization_blend: float = 0.1
    initial_noise_fraction: float = 1e-2
    hyper_shape: float = 0.0
    hyper_rate: float = 0.0
    noise_shape: float = 0.0
    noise_rate: float = 0.0
    noise_update: str = "em"
    max_iter: int = 200
    min_iter: int = 1
    uamp_min_iter: int = 3
    uamp_iteration_offset: float = 0.0
    uamp_iteration_scale: float = 0.8
    polish_size: int = 32
    polish_iterations: int = 2
    polish_damping: float = 0.0
    exchange_size: int = 0
    margin_factor: float = 64.0
    fixed_budget_diagnostics: bool = False
    false_alarm_budget: float | None = 0.5
    jitter_multiplier: float = 8.0
    prune_tail: bool = True
    reinitialize: bool = True
    evidence_gate: bool = True
    truncation_diagnostic: bool = False
    tol: float = 0.0
    uamp_auto_iterations: bool = True
    fixed_noise_variance: float | None = None
    convergence_check_interval: int = 2
    convergence_patience: int = 2

    @staticmethod
    def _adaptive(damping, n, m, max_damping=0.8):
        """uamp_adaptive_damping (solver.py:1180-1186)."""
        return max(damping, min(max_damping, 0.02 * n / m))

    @classmethod
    def fast_hybrid(cls, n: int, m: int, **over) -> "PlanConfig":
        """SBLConfig.fast_hybrid_uamp (solver.py:198-227) over the SBLConfig
        defaults (solver.py:34-104): the default of make_cuda_graph_sbl_plan."""
        kw = dict(uamp_damping=cls._adaptive(0.1, n, m), hyper_shape=1e-8, hyper_rate=1e-8,
                  noise_shape=1e-8, noise_rate=1e-8, min_iter=1, uamp_min_iter=3,
                  uamp_iteration_offset=1.0, uamp_iteration_scale=1.0, polish_iterations=2,
                  margin_factor=0.0, false_alarm_budget=None, truncation_diagnostic=True)
        kw.update(over)
        return cls(**kw)

    @classmethod
    def fast_uamp(cls, n: int, m: int, **over) -> "PlanConfig":
        """SBLConfig.fast_uamp (solver.py:178-195): one reduced evaluation
        with diffuse tail loading, no pruning, re-initialisation or gate."""
        kw = dict(uamp_damping=cls._adaptive(0.1, n, m), hyper_shape=1e-8, hyper_rate=1e-8,
                  noise_shape=1e-8, noise_rate=1e-8, min_iter=3, uamp_min_iter=10,
                  uamp_iteration_offset=10.0, uamp_iteration_scale=1.1, polish_iterations=1,
                  margin_factor=0.0, false_alarm_budget=None, prune_tail=False,
                  reinitialize=False, evidence_gate=False)
        kw.update(over)
        return cls(**kw)

    @classmethod
    def uamp_default(cls, n: int, m: int, **over) -> "PlanConfig":
        """SBLConfig(mode='uamp') with every other field at its default
        (solver.py:34-104): adaptive UAMP damping, T_u = ceil(10 + 1.1 N / M)
        with at least 10 iterations, convergence checks at tol 1e-4 every 2
        iterations with patience 2, and no polish (the UAMP posterior is the
        result)."""
        kw = dict(uamp_damping=cls._adaptive(0.1, n, m), hyper_shape=1e-8, hyper_rate=1e-8,
                  noise_shape=1e-8, noise_rate=1e-8, min_iter=3, uamp_min_iter=10,
                  uamp_iteration_offset=10.0, uamp_iteration_scale=1.1, polish_size=0,
                  polish_iterations=1, margin_factor=0.0, false_alarm_budget=None, prune_tail=False,
                  reinitialize=False, evidence_gate=False, tol=1e-4)
        kw.update(over)
        return cls(**kw)

    @classmethod
    def fast(cls, polish_size: int = 32) -> "PlanConfig":
        return cls(polish_size=polish_size, polish_iterations=2,
                   exchange_size=0, fixed_budget_diagnostics=False)

    @classmethod
    def stable(cls, polish_size: int = 32) -> "PlanConfig":
        return cls(polish_size=polish_size, polish_iterations=3,
                   exchange_size=12, fixed_budget_diagnostics=True)


def uamp_iteration_budget(n: int, m: int, cfg: PlanConfig) -> int:
    """T_u = min(max_iter, max(max(min_iter, uamp_min_iter),
    ceil(offset + scale N/M)))  (solver.py:1280-1309); max_iter without
    uamp_auto_iterations (solver.py:1294-1295)."""
    if not cfg.uamp_auto_iterations:
        return cfg.max_iter
    minimum = max(cfg.min_iter, cfg.uamp_min_iter)
    budget = ceil(cfg.uamp_iteration_offset
                  + cfg.uamp_iteration_scale * n / m)
    return min(cfg.max_iter, max(minimum, budget))


# --------------------------------------------------------------------------
# FFT-UAMP (solver.py:1063-1189, 1390-1925, 2726-2870; triton_kernels.py)
# --------------------------------------------------------------------------

@dataclass
class UampTrace:
    gamma: list = field(default_factory=list)
    noise: list = field(default_factory=list)
    dof: list = field(default_factory=list)


def uamp_initial_state(y, grid, samples, cfg: PlanConfig, eps, tiny, mask=None):
    """Backprojection initialisation (solver.py:1063-1189)."""
    n = grid[0] * grid[1]
    m = y.shape[1]
    power = np.maximum(np.mean(np.abs(y) ** 2, axis=1), eps)
    floor = np.maximum(power * cfg.min_variance, tiny)
    bp = masked_adjoint(y, grid, samples, mask)
    energy = np.abs(bp) ** 2
    structured = (energy / np.maximum(energy.sum(axis=1, keepdims=True), eps)
                  * power[:, None])
    gamma = (cfg.initialization_blend * structured
             + (1.0 - cfg.initialization_blend) * (power[:, None] / n))
    gamma = np.maximum(gamma, floor[:, None])
    noise = np.maximum(cfg.initial_noise_fraction * power, floor)
    if cfg.fixed_noise_variance is not None:
        noise = np.full_like(power, cfg.fixed_noise_variance)      # solver.py:1151-1156
    state = dict(
        mean=np.zeros_like(bp), variance=gamma.copy(),
        message=np.zeros((y.shape[0], m), np.complex128),
        prediction=np.zeros((y.shape[0], m), np.complex128),
        data_variance=gamma.sum(axis=1),
    )
    return power, floor, bp, gamma, noise, state


def uamp_step(y, gamma, noise, state, grid, samples, *, damping, tiny,
              backprojection=None, prior_floor=None, cfg=None,
              compute_prediction=True, mask=None):
    """One UAMP step: message (triton_kernels.py:36-89), lift (A^H or the
    scaled backprojection at the first step, solver.py:2775-2782), denoise +
    optional fused ARD update (triton_kernels.py:92-192), prediction A mu
    (solver.py:2856-2860)."""
    m = y.shape[1]
    v_old = np.maximum(state["variance"], tiny)
    dv = np.maximum(state["data_variance"], tiny)
    total = dv + noise
    precision = 1.0 / total
    pseudo = state["prediction"] - dv[:, None] * state["message"]
    candidate = precision[:, None] * (y - pseudo)
    message = damping * state["message"] + (1.0 - damping) * candidate
    tau = np.maximum(total / m, tiny)
    if backprojection is not None:
        lifted = backprojection * ((1.0 - damping) / total)[:, None]
    else:
        lifted = masked_adjoint(message, grid, samples, mask)
    denom = np.maximum(gamma + tau[:, None], tiny)
    gain = gamma / denom
    cand_mean = gain * (state["mean"] + tau[:, None] * lifted)
    cand_var = gamma * tau[:, None] / denom
    mean = damping * state["mean"] + (1.0 - damping) * cand_mean
    var = np.maximum(damping * v_old + (1.0 - damping) * cand_var, tiny)
    rel = np.clip(1.0 - var / np.maximum(gamma, tiny), 0.0, 1.0)
    dof = rel.sum(axis=1)
    next_prior = None
    if prior_floor is not None:
        second = np.abs(mean) ** 2 + var
        cand_prior = (second + cfg.hyper_rate) / (1.0 + cfg.hyper_shape)
        cand_prior = np.maximum(prior_floor[:, None],
                                np.minimum(cfg.max_precision, cand_prior))
        nxt = cfg.damping * gamma + (1.0 - cfg.damping) * cand_prior
        next_prior = np.maximum(nxt, prior_floor[:, None])
    prediction = (masked_forward(mean, grid, samples, mask) if compute_prediction
                  else state["prediction"])
    new_state = dict(mean=mean, variance=var, message=message,
                     prediction=prediction, data_variance=var.sum(axis=1))
    return mean, var, dof, next_prior, new_state


def uamp_sweep(y, grid, samples, cfg: PlanConfig, *, eps=F32_EPS,
               tiny=F32_TINY, trace: UampTrace | None = None, mask=None):
    """Init + T_u UAMP iterations + final E-step (solver.py:1063-1925);
    y holds the observed samples."""
    n = grid[0] * grid[1]
    m = y.shape[1]
    power, floor, bp, gamma, noise, state = uamp_initial_state(
        y, grid, samples, cfg, eps, tiny, mask)
    t_u = uamp_iteration_budget(n, m, cfg)
    max_dof = m * (1.0 - 10.0 * eps)
    # convergence checks (solver.py:1285-1312, 1800-1904): from the larger
    # of min_iter and uamp_min_iter, every convergence_check_interval
    # iterations and at the budget, on the batch's largest relative change
    # of the UAMP mean; stop after `patience` checks in a row within tol
    minimum = max(cfg.min_iter, cfg.uamp_min_iter)
    convergence = cfg.tol > 0.0 and minimum < t_u
    previous = np.zeros_like(bp)
    streak, converged, change, done = 0, False, float("inf"), t_u
    for it in range(1, t_u + 1):
        mean_t, _, dof, next_prior, state = uamp_step(
            y, gamma, noise, state, grid, samples, damping=cfg.uamp_damping,
            tiny=tiny, backprojection=bp if it == 1 else None,
            prior_floor=floor, cfg=cfg, mask=mask)
        change = float(np.max(np.linalg.norm(mean_t - previous, axis=1)
                              / np.maximum(np.linalg.norm(mean_t, axis=1), eps)))
        previous = mean_t
        dof = np.minimum(dof, max_dof)
        energy = np.sum(np.abs(y - state["prediction"]) ** 2, axis=1)
        if cfg.noise_update == "em":
            cand = ((energy + noise * dof + cfg.noise_rate)
                    / (m + cfg.noise_shape))
        else:
            cand = ((energy + cfg.noise_rate)
                    / np.maximum(m - dof + cfg.noise_shape, eps))
        cand = np.maximum(cand, floor)
        noise = np.maximum(cfg.damping * noise + (1.0 - cfg.damping) * cand,
                           floor)
        if cfg.fixed_noise_variance is not None:
            noise = np.maximum(cfg.fixed_noise_variance, floor)      # solver.py:1747-1750, 1788-1798
        gamma = np.maximum(next_prior, floor[:, None])
        if trace is not None:
            trace.gamma.append(gamma.copy())
            trace.noise.append(noise.copy())
            trace.dof.append(dof.copy())
        check = convergence and it >= minimum and (
            (it - minimum) % cfg.convergence_check_interval == 0 or it == t_u)
        if check:
            streak = streak + 1 if change <= cfg.tol else 0
            if streak >= cfg.convergence_patience:
                converged, done = True, it
                break
    mean, var, dof, _, state = uamp_step(
        y, gamma, noise, state, grid, samples, damping=cfg.uamp_damping,
        tiny=tiny, compute_prediction=False, mask=mask)
    return dict(mean=mean, variance=var, gamma=gamma, noise=noise,
                floor=floor, power=power, backprojection=bp, iterations=done,
                converged=converged, relative_change=change,
                dof=np.minimum(dof, max_dof))


# --------------------------------------------------------------------------
# Exact reduced posterior with evidence (solver.py:3591-3779, 2873-2931)
# --------------------------------------------------------------------------

def reduced_posterior(gram, column_data, data_energy, gamma_s, nu, m, *,
                      eps, tiny, jitter_multiplier=8.0):
    """Equilibrated Cholesky posterior and -log evidence of the zero-tail
    model (solver.py:3591-3779).  Shapes: gram [B,K,K], column_data [B,K],
    gamma_s [B,K], nu [B].  Returns mu_s, sigma_s, data_diag_s, nle."""
    b, k = gamma_s.shape
    system = np.zeros((b, k, k), np.complex128)
    idx = np.arange(k)
    system[:, idx, idx] = 1.0 / gamma_s
    system = system + gram / nu[:, None, None]
    system = 0.5 * (system + np.conj(np.swapaxes(system, 1, 2)))
    dscale = np.sqrt(np.maximum(np.real(system[:, idx, idx]), tiny))
    equil = system / dscale[:, :, None] / dscale[:, None, :]
    equil = 0.5 * (equil + np.conj(np.swapaxes(equil, 1, 2)))
    # GPU branch of _stable_cholesky_raw: one deterministic loading
    # jitter_multiplier * eps * mean|diag| (solver.py:2907-2920).
    scale = np.maximum(np.mean(np.abs(np.real(equil[:, idx, idx])), axis=1),
                       tiny)
    loaded = equil + (jitter_multiplier * eps * scale)[:, None, None] * np.eye(k)
    chol = np.linalg.cholesky(loaded)
    rhs = column_data / nu[:, None]
    scaled = rhs / dscale
    z = np.stack([np.linalg.solve(chol[i], scaled[i]) for i in range(b)])
    emean = np.stack([np.linalg.solve(np.conj(chol[i].T), z[i])
                      for i in range(b)])
    linv = np.stack([np.linalg.solve(chol[i], np.eye(k)) for i in range(b)])
    evar = np.sum(np.abs(linv) ** 2, axis=1)
    mu_s = emean / dscale
    sigma_s = np.maximum(evar / dscale**2, 0.0)
    data_diag = np.maximum((gamma_s - sigma_s) / gamma_s**2, 0.0)
    logdet_sys = 2.0 * (np.sum(np.log(np.maximum(np.real(
        chol[:, idx, idx]), tiny)), axis=1) + np.sum(np.log(dscale), axis=1))
    logdet = m * np.log(nu) + np.sum(np.log(gamma_s), axis=1) + logdet_sys
    quad = np.maximum(data_energy / nu
                      - np.real(np.sum(np.conj(rhs) * mu_s, axis=1)), 0.0)
    return mu_s, sigma_s, data_diag, logdet + quad


def active_deletion_gain(mu_s, sigma_s, gamma_s, tiny):
    """Sherman-Morrison leave-one-out deletion gain (solver.py:3219-3317)."""
    g = np.maximum(gamma_s, tiny)
    q = mu_s / g
    s = np.maximum((g - sigma_s) / g**2, tiny)
    down = np.clip(sigma_s / g, tiny, 1.0)
    return -np.log(down) - g * np.abs(q) ** 2 / down


# --------------------------------------------------------------------------
# Active-set EM polish (solver.py:3961-4511)
# --------------------------------------------------------------------------

def _topk_desc(values, k):
    """Indices of the k largest values per row, ordered by decreasing value
    (torch.topk(sorted=True) order)."""
    order = np.argsort(-values, axis=1, kind="stable")
    return order[:, :k]


def polish(y, sweep, grid, samples, cfg: PlanConfig, *, eps=F32_EPS,
           tiny=F32_TINY, gram_fn=None, mask=None):
    """Evidence-gated active-set EM polish with optional residual exchange;
    y holds the observed samples."""
    b, m = y.shape
    n = grid[0] * grid[1]
    if gram_fn is None:
        def gram_fn(sup, grid_, samples_):
            return masked_gram(sup, grid_, samples_, mask)
    k = min(cfg.polish_size, n)
    floor = sweep["floor"]
    score = np.abs(sweep["mean"]) ** 2 + sweep["variance"]
    support = _topk_desc(score, k)
    rows = np.arange(b)[:, None]
    # solver.py:4087-4106: zero-tail pruning to the floor, re-initialisation
    # of gamma_S from the UAMP second moments
    if cfg.prune_tail:
        working = np.broadcast_to(floor[:, None], (b, n)).copy()
        working[rows, support] = sweep["gamma"][rows, support]
    else:
        working = sweep["gamma"].copy()
    if cfg.reinitialize:
        working[rows, support] = np.maximum(score[rows, support], floor[:, None])
    bp = sweep["backprojection"]
    data_energy = np.sum(np.abs(y) ** 2, axis=1)
    noise = sweep["noise"].copy()

    def tail_of(prior, sup):
        """Diffuse tail loading sum(gamma) - sum(gamma_S) (solver.py:3627-3633,
        4342-4350); zero for the pruned model."""
        if cfg.prune_tail:
            return np.zeros(b)
        return np.maximum(prior.sum(1) - prior[rows, sup].sum(1), 0.0)

    def evaluate(prior, nu, sup):
        gram = gram_fn(sup, grid, samples)
        cdat = bp[rows, sup]          # C^H y = (A^H y)[S]
        mu_s, sig_s, dd_s, nle = reduced_posterior(
            gram, cdat, data_energy, prior[rows, sup], nu + tail_of(prior, sup), m, eps=eps,
            tiny=tiny, jitter_multiplier=cfg.jitter_multiplier)
        mean = np.zeros((b, n), np.complex128)
        mean[rows, sup] = mu_s
        pvar = prior.copy()
        pvar[rows, sup] = sig_s
        ddiag = np.zeros((b, n))
        ddiag[rows, sup] = dd_s
        return dict(mean=mean, pvar=pvar, ddiag=ddiag, nle=nle, prior=prior,
                    noise=nu.copy(), support=sup.copy(), gram=gram)

    best = None
    trace = []
    exchange_pending = None
    exchange_accepted = (np.zeros(b, np.int64) if cfg.exchange_size > 0
                         else None)
    for pit in range(cfg.polish_iterations):
        cand = evaluate(working, noise, support)
        if best is None or not cfg.evidence_gate:
            best = cand
        else:
            margin = cfg.margin_factor * eps * (
                1.0 + np.abs(cand["nle"]) + np.abs(best["nle"]))
            accept = np.isfinite(cand["nle"]) & (
                ~np.isfinite(best["nle"]) | (cand["nle"] <= best["nle"] - margin))
            for key in ("mean", "pvar", "ddiag", "prior", "support", "gram"):
                best[key] = np.where(
                    accept.reshape((b,) + (1,) * (best[key].ndim - 1)),
                    cand[key], best[key])
            best["noise"] = np.where(accept, cand["noise"], best["noise"])
            best["nle"] = np.where(accept, cand["nle"], best["nle"])
            if exchange_pending is not None and exchange_accepted is not None:
                exchange_accepted = exchange_accepted + (
                    accept & exchange_pending).astype(np.int64)
                exchange_pending = None
        trace.append(best["nle"].copy())
        if pit + 1 == cfg.polish_iterations:
            break
        support = best["support"].copy()
        base_prior = best["prior"]
        base_noise = best["noise"]
        mean = best["mean"]
        gamma_s = base_prior[rows, support]
        sig_s = best["pvar"][rows, support]
        mu_s = mean[rows, support]
        new_gamma = (np.abs(mu_s) ** 2 + sig_s + cfg.hyper_rate) / (
            1.0 + cfg.hyper_shape)
        new_gamma = np.minimum(np.maximum(new_gamma, floor[:, None]),
                               cfg.max_precision)
        keep = cfg.polish_damping
        upd_gamma = keep * gamma_s + (1.0 - keep) * new_gamma
        gram = best["gram"]
        cdat = bp[rows, support]
        # ||y - C mu||^2 = ||y||^2 - 2 Re(mu^H C^H y) + mu^H G mu
        energy = (data_energy - 2.0 * np.real(np.sum(np.conj(mu_s) * cdat, 1))
                  + np.real(np.einsum("bk,bkl,bl->b", np.conj(mu_s), gram, mu_s)))
        energy = np.maximum(energy, 0.0)
        dof = np.clip(1.0 - sig_s / np.maximum(gamma_s, tiny), 0.0, 1.0).sum(1)
        # EM on nu = noise + tail, then the tail removed (solver.py:4342-4383)
        inactive = tail_of(base_prior, support)
        if cfg.noise_update == "em":
            upd_noise = (energy + (base_noise + inactive) * dof + cfg.noise_rate) / (
                m + cfg.noise_shape) - inactive
        else:
            upd_noise = (energy + cfg.noise_rate) / np.maximum(
                m - dof + cfg.noise_shape, eps) - inactive
        upd_noise = np.maximum(upd_noise, floor)
        noise = keep * base_noise + (1.0 - keep) * upd_noise
        if cfg.fixed_noise_variance is not None:
            noise = np.maximum(cfg.fixed_noise_variance, floor)      # solver.py:4358-4387
        working = base_prior.copy()
        working[rows, support] = upd_gamma
        exch = min(cfg.exchange_size, max(k - 1, 0), n - k)
        if exch > 0 and pit == 0:
            # |A^H r|^2 with r = y - C mu_S; A^H C mu_S via the Dirichlet
            # kernel: (A^H A)[n, k] = D1[n1-k1] D2[n2-k2].
            resid_bp = bp - _gram_apply(mu_s, support, grid, samples, mask)
            rscore = np.abs(resid_bp) ** 2
            rscore[rows, support] = -np.inf
            add_idx = _topk_desc(rscore, exch)
            add_energy = rscore[rows, add_idx]
            dgain = active_deletion_gain(mu_s, sig_s, gamma_s, tiny)
            rem_pos = _topk_desc(dgain, exch)
            rem_idx = support[rows, rem_pos]
            col_energy = float(m)
            ratio = add_energy / np.maximum(base_noise[:, None] * col_energy,
                                            tiny)
            add_gain = np.where(ratio > 1.0,
                                ratio - 1.0 - np.log(np.maximum(ratio, tiny)),
                                0.0)
            del_gain = dgain[rows, rem_pos]
            swap = (add_gain + del_gain) > 0.0
            repl_idx = np.where(swap, add_idx, rem_idx)
            new_support = support.copy()
            new_support[rows, rem_pos] = repl_idx
            add_gamma = (np.maximum(add_energy - base_noise[:, None]
                                    * col_energy, 0.0) / col_energy**2)
            add_gamma = np.minimum(np.maximum(add_gamma, floor[:, None]),
                                   cfg.max_precision)
            retained = working[rows, rem_idx]
            working[rows, rem_idx] = np.where(swap, floor[:, None], retained)
            working[rows, repl_idx] = np.where(swap, add_gamma, retained)
            support = new_support
            exchange_pending = swap.any(axis=1)
    out = dict(
        mean=best["mean"], posterior_variance=best["pvar"],
        data_diagonal=best["ddiag"], gamma=best["prior"],
        noise=best["noise"], negative_log_evidence=best["nle"],
        evidence_trace=np.stack(trace, axis=1), support=best["support"],
        exchanges_accepted=exchange_accepted, initial_support=_topk_desc(
            score, k),
    )
    if cfg.truncation_diagnostic and cfg.prune_tail:
        out["truncation_delta_bound"] = truncation_delta_bound(
            sweep["gamma"], sweep["noise"], out["initial_support"], grid, samples, tiny)
    return out


def smooth_length(minimum: int) -> int:
    """First 2/3/5/7-smooth length >= minimum (operators.py:10-20)."""
    c = int(minimum)
    while True:
        r = c
        for f in (2, 3, 5, 7):
            while r % f == 0:
                r //= f
        if r == 1:
            return c
        c += 1


def truncation_delta_bound(gamma, noise, support, grid, samples, tiny=F32_TINY):
    """Hard-pruning certificate (solver.py:3927-3958): the largest eigenvalue
    magnitude of the circulant embedding of the tail covariance
    A Gamma_tail A^H (gamma with the support zeroed), over the noise.  The
    embedding is E1 x E2 = smooth(2M - 1) per axis with the lag d at
    position d mod E (operators.py:100-130, 299-326)."""
    b = gamma.shape[0]
    n1, n2 = grid
    m1, m2 = samples
    tail = np.array(gamma, np.float64).copy()
    tail[np.arange(b)[:, None], support] = 0.0
    lags = np.fft.fft2(tail.reshape(b, n1, n2), axes=(1, 2))
    e1, e2 = smooth_length(2 * m1 - 1), smooth_length(2 * m2 - 1)
    d1 = np.arange(-(m1 - 1), m1)
    d2 = np.arange(-(m2 - 1), m2)
    kern = np.zeros((b, e1, e2), np.complex128)
    kern[:, (d1 % e1)[:, None], (d2 % e2)[None, :]] = lags[:, (d1 % n1)[:, None], (d2 % n2)[None, :]]
    spec = np.fft.fft2(kern, axes=(1, 2))
    return np.abs(spec).reshape(b, -1).max(1) / np.maximum(noise, tiny)


def _gram_apply(mu_s, support, grid, samples, mask=None):
    """(A^H A x_S)[n] = sum_k mu_k D1[(n1-k1) mod N1] D2[(n2-k2) mod N2];
    with a mask, A^H (A x_S) on the observed rows through the transforms."""
    if mask is not None:
        cols = restrict(fourier_columns(support, grid, samples), mask)
        return masked_adjoint(np.einsum("bmk,bk->bm", cols, mu_s), grid, samples, mask)
    n1, n2 = grid
    d1 = dirichlet(samples[0], n1)
    d2 = dirichlet(samples[1], n2)
    k1 = support // n2
    k2 = support % n2
    r1 = np.arange(n1)
    r2 = np.arange(n2)
    out = np.zeros((support.shape[0], n1, n2), np.complex128)
    for bi in range(support.shape[0]):
        for j in range(support.shape[1]):
            a = d1[(r1 - k1[bi, j]) % n1]
            c = d2[(r2 - k2[bi, j]) % n2]
            out[bi] += mu_s[bi, j] * np.outer(a, c)
    return out.reshape(support.shape[0], -1)


def fixed_budget_residuals(y, result, floor, grid, samples, cfg, *, eps,
                           tiny, mask=None):
    """Stable's EM fixed-point residuals (solver.py:3408-3496) and the
    evidence-trace monotonicity flags (solver.py:1989-2021)."""
    b, m = y.shape
    rows = np.arange(b)[:, None]
    sup = result["support"]
    gamma_s = np.maximum(result["gamma"][rows, sup], tiny)
    sig_s = result["posterior_variance"][rows, sup]
    mu_s = result["mean"][rows, sup]
    target = (np.abs(mu_s) ** 2 + sig_s + cfg.hyper_rate) / (1.0 + cfg.hyper_shape)
    target = np.minimum(np.maximum(target, floor[:, None]), cfg.max_precision)
    scale = np.maximum(np.maximum(np.abs(gamma_s), np.abs(target)), tiny)
    gamma_res = np.max(np.abs(gamma_s - target) / scale, axis=1)
    dof = np.clip(1.0 - sig_s / gamma_s, 0.0, 1.0).sum(1)
    cols = restrict(fourier_columns(sup, grid, samples), mask)
    pred = np.einsum("bmk,bk->bm", cols, mu_s)
    energy = np.sum(np.abs(y - pred) ** 2, axis=1)
    noise = result["noise"]
    if cfg.noise_update == "em":
        ntarget = (energy + noise * dof + cfg.noise_rate) / (m + cfg.noise_shape)
    else:
        ntarget = (energy + cfg.noise_rate) / np.maximum(
            m - dof + cfg.noise_shape, eps)
    ntarget = np.minimum(np.maximum(ntarget, floor), cfg.max_precision)
    nscale = np.maximum(np.maximum(np.abs(noise), np.abs(ntarget)), tiny)
    noise_res = np.abs(noise - ntarget) / nscale
    if cfg.fixed_noise_variance is not None:
        noise_res = np.zeros_like(noise_res)                          # solver.py:3457-3458
    tr = result["evidence_trace"]
    finite = np.all(np.isfinite(tr), axis=1)
    nonincreasing = np.all(tr[:, 1:] <= tr[:, :-1], axis=1)
    return dict(gamma_residual=gamma_res, noise_residual=noise_res,
                monotone=finite & nonincreasing,
                objective_decrease=np.maximum(tr[:, 0] - tr[:, -1], 0.0))


def project_relevance_dof(gamma, data_diag, m, eps):
    """sum of projected relevance (solver.py:3156-3181)."""
    rel = np.clip(gamma * data_diag, 0.0, 1.0)
    raw = rel.sum(axis=1)
    scale = np.minimum(1.0, m * (1.0 - 10.0 * eps) / np.maximum(raw, eps))
    return (rel * scale[:, None]).sum(axis=1)


def fbme_plan(data: np.ndarray, grid, profile: str = "fast",
              polish_size: int = 32, *, eps=F32_EPS, tiny=F32_TINY,
              mask=None, config: PlanConfig | None = None) -> dict:
    """End-to-end Fast/Stable plan on [B, M1, M2] (or [M1, M2]) data,
    optionally on the samples an observation mask [M1, M2] keeps.

    Returns the SBLResult fields the locked plans populate
    (solver.py:2081-2126, 2154-2393)."""
    data = np.asarray(data)
    single = data.ndim == 2
    if single:
        data = data[None]
    b = data.shape[0]
    samples = tuple(int(v) for v in data.shape[-2:])
    grid = tuple(int(v) for v in grid)
    n = grid[0] * grid[1]
    if mask is not None:
        mask = np.asarray(mask, bool)
        if mask.shape != samples or not mask.any():
            raise ValueError("observation_mask must match the samples and keep one")
    y = restrict(data.reshape(b, -1).astype(np.complex128), mask)
    m = y.shape[1]
    cfg = config if config is not None else (
        PlanConfig.fast if profile == "fast" else PlanConfig.stable)(polish_size)
    sweep = uamp_sweep(y, grid, samples, cfg, eps=eps, tiny=tiny, mask=mask)
    if cfg.polish_size == 0:
        # the UAMP posterior of the final hyperparameters (solver.py:1907-1925,
        # 2081-2087)
        out = dict(x=sweep["mean"], posterior_variance=sweep["variance"],
                   precision=1.0 / sweep["gamma"], noise_variance=sweep["noise"],
                   effective_dof=sweep["dof"], iterations=sweep["iterations"],
                   converged=sweep["converged"], relative_change=sweep["relative_change"],
                   gamma=sweep["gamma"], floor=sweep["floor"])
        if single:
            out = {k: (v[0] if isinstance(v, np.ndarray) and v.shape[:1] == (1,) else v)
                   for k, v in out.items()}
        return out
    res = polish(y, sweep, grid, samples, cfg, eps=eps, tiny=tiny, mask=mask)
    out = dict(
        x=res["mean"], posterior_variance=res["posterior_variance"],
        precision=1.0 / res["gamma"], noise_variance=res["noise"],
        effective_dof=project_relevance_dof(res["gamma"], res["data_diagonal"],
                                            m, eps),
        iterations=sweep["iterations"],
        negative_log_evidence=res["negative_log_evidence"],
        evidence_trace=res["evidence_trace"], support=res["support"],
        initial_support=res["initial_support"], gamma=res["gamma"],
        uamp_noise=sweep["noise"], uamp_mean=sweep["mean"],
        uamp_variance=sweep["variance"], uamp_gamma=sweep["gamma"],
        floor=sweep["floor"],
        support_exchanges_accepted=res["exchanges_accepted"],
    )
    if "truncation_delta_bound" in res:
        out["truncation_delta_bound"] = res["truncation_delta_bound"]
    if cfg.fixed_budget_diagnostics:
        out.update({f"fixed_budget_{k}": v for k, v in fixed_budget_residuals(
            y, res, sweep["floor"], grid, samples, cfg, eps=eps,
            tiny=tiny, mask=mask).items()})
    if cfg.false_alarm_budget is not None:
        thr = log(n / cfg.false_alarm_budget)
        score = np.abs(out["x"]) ** 2 / np.maximum(out["posterior_variance"],
                                                   tiny)
        out["support_mask"] = score > thr
        out["support_score_threshold"] = np.full(b, thr)
    if single:
        out = {k: (v[0] if isinstance(v, np.ndarray) and v.shape[:1] == (1,)
                   else v) for k, v in out.items()}
    return out
