"""Mapping from SBLConfig to the native FBME plan and back to SBLResult.

The native plan implements the UAMP + evidence-gated active-set EM path of
fast_sbl (torch_sbl/solver.py:925-2393) for the settings the locked plans and
the fast/paper hybrid presets use.  Options that select research paths of
the reference outside that path raise NotImplementedError naming the option,
so nothing silently runs a different algorithm.
"""

from __future__ import annotations

import math
from collections import OrderedDict
from math import prod

import torch

from fbme_amd import _lib, autotune
from fbme_amd.plan import NativePlan

from .config import SBLConfig, SBLResult

_PLAN_CACHE: "OrderedDict[tuple, NativePlan]" = OrderedDict()
_PLAN_CACHE_SIZE = 4


def clear_cache() -> None:
    _PLAN_CACHE.clear()


def _unsupported(cfg: SBLConfig) -> list[str]:
    bad = []
    checks = {
        "uamp_polish_residual_exchange_size>0 without prune_tail and evidence_gate": (
            cfg.uamp_polish_residual_exchange_size > 0
            and not (cfg.uamp_polish_prune_tail and cfg.uamp_polish_evidence_gate)),
        "uamp_polish_fixed_budget_diagnostics without prune_tail": (
            cfg.uamp_polish_fixed_budget_diagnostics and not cfg.uamp_polish_prune_tail),
        "uamp_polish_global_certificate": cfg.uamp_polish_global_certificate,
        "uamp_polish_support_repairs": cfg.uamp_polish_support_repairs != 0,
        "uamp_polish_support_growth_steps": (
            cfg.uamp_polish_support_growth_steps != 0),
        "uamp_polish_support_shrink_steps": (
            cfg.uamp_polish_support_shrink_steps != 0),
        "uamp_polish_support_prior": cfg.uamp_polish_support_prior,
        "uamp_polish_high_precision": cfg.uamp_polish_high_precision,
        "uamp_polish_coordinate_steps": cfg.uamp_polish_coordinate_steps != 0,
        "uamp_polish_noise_coordinate_steps": (
            cfg.uamp_polish_noise_coordinate_steps != 0),
        "joint_sparsity": cfg.joint_sparsity,
        "spatial coupling": (cfg.spatial_coupling > 0.0
                             or cfg.learn_spatial_coupling),
        "phase_error_correction": cfg.phase_error_correction,
        "range_migration_correction": cfg.range_migration_correction,
        "learn_snapshot_covariance": cfg.learn_snapshot_covariance,
        f"uamp_polish_size>{32}": cfg.uamp_polish_size > 32,
    }
    for name, flag in checks.items():
        if flag:
            bad.append(name)
    return bad


def check_supported(cfg: SBLConfig) -> None:
    bad = _unsupported(cfg)
    if bad:
        raise NotImplementedError(
            "The MI355X FBME plan implements the UAMP + evidence-gated "
            "active-set path of the fast/paper hybrid presets; unsupported "
            "options: " + ", ".join(bad))


def plan_options(cfg: SBLConfig, grid_shape, sample_shape, observed: int | None = None
                 ) -> _lib.FbmePlanOptions:
    """FbmePlanOptions equivalent of a UAMP-mode SBLConfig (`observed`: the
    sample count an observation mask keeps)."""
    n = prod(grid_shape)
    m = prod(sample_shape) if observed is None else int(observed)
    damping = cfg.uamp_damping
    if cfg.uamp_adaptive_damping:
        # solver.py:1180-1186
        damping = max(damping, min(cfg.uamp_max_damping, 0.02 * n / m))
    opt = _lib.default_options(cfg.uamp_polish_residual_exchange_size > 0)
    opt.uamp_damping = float(damping)
    opt.damping = float(cfg.damping)
    opt.min_variance = float(cfg.min_variance)
    opt.max_precision = float(cfg.max_precision)
    # 'uniform' is the blend 0 (solver.py:1070-1089); its first lifted
    # message is then A^H of the first message, the backprojection scaled by
    # the same factor (solver.py:2823-2830)
    opt.blend = float(cfg.initialization_blend) if cfg.initialization == "backprojection" else 0.0
    opt.noise_fraction = float(cfg.initial_noise_fraction)
    opt.margin_factor = float(cfg.uamp_polish_evidence_margin_factor)
    opt.jitter_multiplier = float(cfg.cuda_cholesky_jitter_multiplier)
    opt.false_alarm = float(cfg.uamp_support_false_alarm_budget or 0.0)
    opt.polish_size = int(min(cfg.uamp_polish_size, n))
    opt.polish_iterations = int(cfg.uamp_polish_iterations)
    opt.exchange_size = int(cfg.uamp_polish_residual_exchange_size)
    opt.diagnostics = 1 if cfg.uamp_polish_fixed_budget_diagnostics else 0
    opt.evidence_noise = 1 if cfg.noise_update == "evidence" else 0
    opt.iteration_offset = float(cfg.uamp_iteration_offset)
    opt.iteration_scale = float(cfg.uamp_iteration_scale)
    if not cfg.uamp_auto_iterations:
        # T_u = max_iter (solver.py:1295-1309): ceil(max_iter + 0 N / M)
        opt.iteration_offset = float(cfg.max_iter)
        opt.iteration_scale = 0.0
    opt.min_iters = int(max(cfg.min_iter, cfg.uamp_min_iter))
    opt.max_iters = int(cfg.max_iter)
    opt.hyper_shape = float(cfg.hyper_shape)
    opt.hyper_rate = float(cfg.hyper_rate)
    opt.noise_shape = float(cfg.noise_shape)
    opt.noise_rate = float(cfg.noise_rate)
    opt.polish_damping = float(cfg.uamp_polish_damping)
    opt.prune_tail = int(cfg.uamp_polish_prune_tail)
    opt.reinitialize = int(cfg.uamp_polish_reinitialize)
    opt.evidence_gate = int(cfg.uamp_polish_evidence_gate)
    opt.fixed_noise = float(cfg.fixed_noise_variance or 0.0)
    # solver.py:4034-4040
    opt.truncation_diagnostic = int(cfg.uamp_polish_prune_tail and (
        cfg.uamp_polish_compute_truncation_diagnostic
        or cfg.uamp_polish_max_truncation_delta is not None))
    return opt


def options_key(opt: _lib.FbmePlanOptions) -> tuple:
    return tuple(getattr(opt, name) for name, _ in opt._fields_)


def tuned(shapes, opt: _lib.FbmePlanOptions, device) -> _lib.FbmePlanOptions:
    """opt with the autotuned FFT pass shapes of this workload on this GPU
    (unchanged when the workload was never tuned)."""
    return autotune.tuned_plan_options(shapes, opt.exchange_size > 0, device, opt)


def cached_plan(shapes, opt, device, mask=None) -> NativePlan:
    """Plan of these shapes and options (one observation mask, shared by
    every scene, or None), built once per key."""
    opt = tuned(shapes, opt, device)
    mkey = None
    if mask is not None:
        mask = torch.as_tensor(mask, dtype=torch.bool).cpu()
        mkey = (tuple(mask.shape), mask.numpy().tobytes())
    key = (tuple(shapes), options_key(opt), str(device), mkey)
    plan = _PLAN_CACHE.get(key)
    if plan is None:
        plan = NativePlan(shapes, opt, device, None if mask is None else [mask] * len(shapes))
        _PLAN_CACHE[key] = plan
        while len(_PLAN_CACHE) > _PLAN_CACHE_SIZE:
            _PLAN_CACHE.popitem(last=False)
    else:
        _PLAN_CACHE.move_to_end(key)
    return plan


class ResultViews:
    """Tensors of an SBLResult that alias a plan's outputs, built once so a
    captured graph keeps ownership of every returned tensor."""

    def __init__(self, plan: NativePlan, cfg: SBLConfig, batch_shape,
                 grid_shape) -> None:
        out = plan.out
        dev = plan.device
        bshape = tuple(batch_shape)
        oshape = (*bshape, *grid_shape)
        b = plan.nscenes
        n = prod(grid_shape)
        self.cfg = cfg
        self.plan = plan
        self.batch_shape = bshape
        self.grid_shape = tuple(grid_shape)
        self.x = out.x.view(oshape)
        self.posterior_variance = out.posterior_variance.view(oshape)
        self.precision = out.precision.view(oshape)
        self.noise = out.noise.view(bshape)
        self.dof = out.dof.view(bshape)
        self.nle = out.nle.view(bshape)
        self.trace = out.trace
        self.hard = torch.zeros(b, dtype=torch.bool, device=dev).view(bshape)
        # solver.py:4041-4062: the certificate, and hard pruning accepted
        # when it is within uamp_polish_max_truncation_delta
        self.truncation = (out.truncation.view(bshape)
                           if plan.options.truncation_diagnostic else None)
        self.selected = torch.ones(b, dtype=torch.bool, device=dev).view(bshape)
        self.penalty = torch.zeros(b, dtype=torch.float32, device=dev).view(bshape)
        self.exchanges = (out.exchanges.view(bshape)
                          if cfg.uamp_polish_residual_exchange_size > 0 else None)
        diag = cfg.uamp_polish_fixed_budget_diagnostics
        self.gres = out.diag[:, 0].view(bshape) if diag else None
        self.nres = out.diag[:, 1].view(bshape) if diag else None
        self.decrease = out.diag[:, 2].view(bshape) if diag else None
        self.monotone = (torch.zeros(b, dtype=torch.bool, device=dev).view(bshape)
                         if diag else None)
        fa = cfg.uamp_support_false_alarm_budget
        if fa is not None:
            self.mask = out.support_mask.view(oshape)
            self.threshold = torch.full((b,), math.log(n / fa),
                                        dtype=torch.float32,
                                        device=dev).view(bshape)
        else:
            self.mask = None
            self.threshold = None
        # T_u of the plan's first scene (N / observed samples, solver.py:1296-1309)
        self.iterations = int(plan.scene_info[0][2])
        self.k = int(plan.options.polish_size)
        if self.k == 0:
            # no polish (solver.py:1926): no evidence, certificate or
            # exchange fields
            self.nle = self.trace = self.truncation = self.hard = None
            self.selected = self.penalty = self.exchanges = None
            self.gres = self.nres = self.decrease = self.monotone = None
        self.converged = False
        self.change = float("nan")
        self.history = {}

    def refresh(self) -> None:
        """Device-side derived fields (enqueued on the current stream)."""
        delta = self.cfg.uamp_polish_max_truncation_delta
        if self.truncation is not None and delta is not None:
            torch.le(self.truncation, float(delta), out=self.hard)
        if self.monotone is not None:
            self.monotone.copy_(self.plan.out.diag[:, 3].view(self.batch_shape) > 0.5)

    def result(self, clone: bool = False) -> SBLResult:
        res = self._result()
        if clone:
            for name in res.__slots__:
                value = getattr(res, name)
                if isinstance(value, torch.Tensor):
                    setattr(res, name, value.clone())
        return res

    def _result(self) -> SBLResult:
        return SBLResult(
            x=self.x,
            posterior_variance=self.posterior_variance,
            precision=self.precision,
            noise_variance=self.noise,
            effective_dof=self.dof,
            iterations=self.iterations,
            converged=self.converged,
            relative_change=self.change,
            mode="uamp",
            arithmetic_dtype="complex64",
            uamp_backend_used="hip",
            active_set_size=self.k,
            negative_log_evidence=self.nle,
            evidence_trace=self.trace,
            truncation_delta_bound=self.truncation,
            hard_pruning_accepted=self.hard,
            truncation_certified=self.hard,
            support_repairs_accepted=self.exchanges,
            support_exchanges_accepted=self.exchanges,
            active_model_selected=self.selected,
            support_log_odds_penalty=self.penalty,
            penalized_negative_log_evidence=self.nle,
            penalized_evidence_trace=self.trace,
            fixed_budget_gamma_residual=self.gres,
            fixed_budget_noise_residual=self.nres,
            fixed_budget_objective_decrease=self.decrease,
            fixed_budget_monotone=self.monotone,
            support_mask=self.mask,
            support_score_threshold=self.threshold,
            history=self.history,
        )


def run_eager(plan: NativePlan, views: ResultViews, y: torch.Tensor, cfg: SBLConfig) -> None:
    """One solve driven iteration by iteration from the host: the convergence
    checks, history and change tracking of the reference's loop
    (solver.py:1285-1312, 1800-1904) around the plan's phases.  Stopping at
    iteration t finishes every scene there (the final E-step and the polish
    or UAMP posterior of fbme_plan_finish)."""
    from fbme_amd.plan import SC_CHANGE, SC_DOF, SC_NOISE
    limit = plan.max_iters
    minimum = max(cfg.min_iter, cfg.uamp_min_iter)
    convergence = cfg.tol > 0.0 and minimum < limit
    track_all = cfg.return_history or cfg.eager_change_tracking or cfg.verbose
    history = {k: [] for k in ("relative_change", "noise_variance", "effective_dof", "cg_iterations",
                               "hutchinson_probes", "cg_tolerance", "active_set_size",
                               "spatial_vertical_coupling", "spatial_horizontal_coupling")}
    scalars = plan.scalars()
    change = float("inf")
    streak = 0
    converged = False
    plan.begin(y)
    it = 0
    for it in range(1, limit + 1):
        check = convergence and it >= minimum and ((it - minimum) % cfg.convergence_check_interval == 0
                                                   or it == limit)
        need = track_all or check or it == limit
        plan.iterate(y, it, need)
        if need:
            sc = scalars.cpu()
            change = float(sc[:, SC_CHANGE].max())
            if cfg.return_history or cfg.verbose:
                noise = float(sc[:, SC_NOISE].float().mean())
                dof = float(sc[:, SC_DOF].float().mean())
        if cfg.return_history:
            history["relative_change"].append(change)
            history["noise_variance"].append(noise)
            history["effective_dof"].append(dof)
            for key in ("cg_iterations", "hutchinson_probes", "cg_tolerance", "active_set_size"):
                history[key].append(0.0)
            history["spatial_vertical_coupling"].append(float(cfg.spatial_coupling))
            history["spatial_horizontal_coupling"].append(float(cfg.spatial_coupling))
        if cfg.verbose and (it == 1 or it % 10 == 0):
            print(f"[SBL] iter={it:4d} change={change:.3e} noise={noise:.3e} dof={dof:.2f} "
                  f"cg={0:3d} probes={0:2d} active=0")
        if check:
            streak = streak + 1 if change <= cfg.tol else 0
            if streak >= cfg.convergence_patience:
                converged = True
                break
    plan.finish(y)
    views.refresh()
    views.iterations = it
    views.converged = converged
    views.change = (change if (cfg.return_history or cfg.eager_change_tracking or cfg.tol > 0.0 or cfg.verbose)
                    else float("nan"))
    views.history = history if cfg.return_history else {}
