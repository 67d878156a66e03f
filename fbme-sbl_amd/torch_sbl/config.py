"""SBLConfig and SBLResult with the reference's field names and defaults.

Field names, defaults and presets follow torch_sbl/solver.py:34-350 (config)
and :720-792 (result) so that configurations and result consumers written for
the reference run unchanged.  Validation (solver.py:352-717) is restated as a
table of rules; fields that only steer research paths outside the FBME hot
path are still validated so that an invalid configuration fails the same way.
"""

from __future__ import annotations

from dataclasses import dataclass, field, fields
from typing import Any

import torch
from torch import Tensor


@dataclass(slots=True)
class SBLConfig:
    """Numerical and iteration settings (torch_sbl/solver.py:34-176)."""

    max_iter: int = 200
    min_iter: int = 3
    tol: float = 1e-4
    hyper_shape: float = 1e-8
    hyper_rate: float = 1e-8
    noise_shape: float = 1e-8
    noise_rate: float = 1e-8
    initial_noise_fraction: float = 1e-2
    initialization: str = "backprojection"
    initialization_blend: float = 0.1
    fixed_noise_variance: float | None = None
    noise_update: str = "em"
    damping: float = 0.15
    min_variance: float = 1e-12
    max_precision: float = 1e12
    mode: str = "auto"
    dense_threshold: int = 2048
    cg_tolerance: float = 1e-5
    cg_initial_tolerance: float = 1e-2
    cg_max_iter: int = 150
    cg_check_interval: int = 8
    cg_fixed_iterations: int = 0
    cg_matvec: str = "tbt"
    cg_preconditioner: str = "low_rank"
    cg_preconditioner_rank: int = 32
    cg_damping: float = 0.9
    probe_type: str = "rademacher"
    hutchinson_probes: int = 32
    max_hutchinson_probes: int = 128
    initial_hutchinson_probes: int = 16
    auto_probe_scaling: bool = True
    adaptive_cg: bool = True
    adaptive_probes: bool = False
    cg_warm_start: bool = True
    cg_min_iter: int = 20
    convergence_patience: int = 2
    convergence_check_interval: int = 2
    cuda_sync_free_cholesky: bool = True
    cuda_cholesky_jitter_multiplier: float = 8.0
    cache_fourier_operator: bool = True
    dense_joint_solve: bool = True
    eager_change_tracking: bool = False
    uamp_damping: float = 0.1
    uamp_adaptive_damping: bool = True
    uamp_max_damping: float = 0.8
    uamp_min_iter: int = 10
    uamp_auto_iterations: bool = True
    uamp_iteration_offset: float = 10.0
    uamp_iteration_scale: float = 1.1
    uamp_polish_size: int = 0
    uamp_polish_iterations: int = 1
    uamp_polish_prune_tail: bool = False
    uamp_polish_compute_truncation_diagnostic: bool = True
    uamp_polish_reinitialize: bool = False
    uamp_polish_damping: float = 0.0
    uamp_polish_evidence_gate: bool = False
    uamp_polish_evidence_margin_factor: float = 0.0
    uamp_polish_residual_exchange_size: int = 0
    uamp_polish_max_truncation_delta: float | None = None
    uamp_polish_global_certificate: bool = False
    uamp_polish_support_repairs: int = 0
    uamp_polish_support_growth_steps: int = 0
    uamp_polish_support_shrink_steps: int = 0
    uamp_polish_support_shrink_batch_size: int = 1
    uamp_polish_support_prior: bool = False
    uamp_polish_support_prior_strength: float = 1.0
    uamp_polish_high_precision: bool = False
    uamp_polish_coordinate_steps: int = 0
    uamp_polish_coordinate_tolerance: float = 1e-6
    uamp_polish_coordinate_early_stop: bool = True
    uamp_polish_noise_coordinate_steps: int = 0
    uamp_polish_noise_tolerance: float = 1e-4
    uamp_polish_noise_objective_tolerance: float = 1e-10
    uamp_polish_joint_cycles: int = 1
    uamp_polish_familywise_error_rate: float | None = None
    uamp_polish_fixed_budget_diagnostics: bool = False
    uamp_support_false_alarm_budget: float | None = None
    uamp_backend: str = "auto"
    active_set: bool = True
    active_set_start: int = 30
    active_set_check_interval: int = 2
    active_set_min_size: int = 32
    active_set_max_size: int = 256
    active_set_tail_fraction: float = 5e-2
    joint_sparsity: bool = False
    joint_noise_variance: bool = True
    spatial_coupling: float = 0.0
    spatial_boundary: str = "zero"
    learn_spatial_coupling: bool = False
    spatial_coupling_max: float = 1.0
    spatial_coupling_damping: float = 0.5
    spatial_coupling_steps: int = 4
    spatial_coupling_update_interval: int = 2
    spatial_coupling_penalty: float = 1e-3
    spatial_anisotropy_penalty: float = 1e-2
    learn_snapshot_covariance: bool = False
    snapshot_covariance_model: str = "ar1"
    snapshot_covariance_damping: float = 0.5
    snapshot_covariance_shrinkage: float = 0.05
    snapshot_covariance_floor: float = 1e-3
    snapshot_covariance_start: int = 20
    snapshot_covariance_warmup: int = 20
    snapshot_covariance_min_correlation: float = 0.5
    snapshot_covariance_min_snr_db: float = 10.0
    snapshot_covariance_min_undersampling: float = 48.0
    phase_error_correction: bool = False
    phase_error_axis: int = 1
    phase_basis_rank: int = 6
    phase_optimization_steps: int = 80
    phase_learning_rate: float = 0.05
    phase_smoothness: float = 5.0
    phase_evidence_penalty: float = 1.0
    phase_min_evidence_gain_per_parameter: float = 1.0
    phase_slq_probes: int = 8
    phase_slq_steps: int = 20
    range_dependent_phase: bool = False
    phase_range_basis_rank: int = 2
    range_migration_correction: bool = False
    migration_basis_rank: int = 4
    migration_max_shift: float = 4.0
    migration_smoothness: float = 1.0
    motion_phase_amplitude_penalty: float = 0.2
    motion_migration_amplitude_penalty: float = 0.2
    motion_refinement_steps: int = 40
    motion_refinement_learning_rate: float = 0.02
    offgrid_refinement: bool = False
    offgrid_components: int | None = None
    offgrid_max_components: int = 64
    offgrid_relative_threshold: float = 0.1
    offgrid_minimum_separation: float = 1.5
    offgrid_max_iter: int = 15
    offgrid_tolerance: float = 1e-4
    offgrid_max_step: float = 0.35
    offgrid_damping: float = 1e-3
    offgrid_ridge: float = 1e-7
    seed: int = 0
    device: Any = "auto"
    return_history: bool = True
    verbose: bool = False

    # ------------------------------------------------------------ presets
    @classmethod
    def fast_uamp(cls, **overrides: Any) -> "SBLConfig":
        """Low-latency FFT-UAMP preset (solver.py:178-196)."""
        settings: dict[str, Any] = dict(
            mode="uamp", tol=0.0, return_history=False,
            uamp_auto_iterations=True, uamp_polish_size=32, active_set=False)
        settings.update(overrides)
        return cls(**settings)

    @classmethod
    def fast_hybrid_uamp(cls, **overrides: Any) -> "SBLConfig":
        """Few-step UAMP plus exact active-set EM (solver.py:198-227)."""
        settings: dict[str, Any] = dict(
            mode="uamp", max_iter=200, min_iter=1, tol=0.0,
            return_history=False, uamp_auto_iterations=True,
            uamp_min_iter=3, uamp_iteration_offset=1.0,
            uamp_iteration_scale=1.0, uamp_polish_size=32,
            uamp_polish_iterations=2, uamp_polish_prune_tail=True,
            uamp_polish_reinitialize=True, uamp_polish_damping=0.0,
            uamp_polish_evidence_gate=True,
            uamp_polish_residual_exchange_size=0, active_set=False)
        settings.update(overrides)
        return cls(**settings)

    @classmethod
    def paper_hybrid_uamp(cls, **overrides: Any) -> "SBLConfig":
        """Fixed-budget monotone-evidence FBME method (solver.py:230-280)."""
        settings: dict[str, Any] = dict(
            device="cuda", uamp_backend="triton", tol=0.0,
            return_history=False, eager_change_tracking=False,
            hyper_shape=0.0, hyper_rate=0.0, noise_shape=0.0, noise_rate=0.0,
            uamp_iteration_offset=0.0, uamp_iteration_scale=0.8,
            uamp_damping=0.05, uamp_adaptive_damping=False,
            uamp_polish_iterations=3, uamp_polish_global_certificate=False,
            uamp_polish_support_repairs=0,
            uamp_polish_support_growth_steps=0,
            uamp_polish_support_shrink_steps=0,
            uamp_polish_support_prior=False,
            uamp_polish_high_precision=False,
            uamp_polish_coordinate_steps=0,
            uamp_polish_noise_coordinate_steps=0,
            uamp_polish_residual_exchange_size=12,
            uamp_polish_evidence_margin_factor=64.0,
            uamp_polish_compute_truncation_diagnostic=False,
            uamp_polish_familywise_error_rate=None,
            uamp_polish_fixed_budget_diagnostics=True,
            uamp_support_false_alarm_budget=0.5,
            phase_error_correction=False, range_migration_correction=False,
            offgrid_refinement=False, learn_snapshot_covariance=False)
        settings.update(overrides)
        return cls.fast_hybrid_uamp(**settings)

    @classmethod
    def realtime_hybrid_uamp(cls, **overrides: Any) -> "SBLConfig":
        """Graph-capturable real-time preset (solver.py:283-298)."""
        return cls.paper_hybrid_uamp(**overrides)

    @classmethod
    def certified_hybrid_uamp(cls, **overrides: Any) -> "SBLConfig":
        """Hybrid preset with Type-II coordinate diagnostics
        (solver.py:300-318)."""
        settings: dict[str, Any] = dict(
            uamp_polish_global_certificate=True, uamp_polish_iterations=4,
            uamp_polish_support_repairs=1)
        settings.update(overrides)
        return cls.fast_hybrid_uamp(**settings)

    @classmethod
    def taes_hybrid_uamp(cls, **overrides: Any) -> "SBLConfig":
        """Evidence-monotone research preset (solver.py:320-350)."""
        settings: dict[str, Any] = dict(
            uamp_polish_global_certificate=True, uamp_polish_iterations=8,
            uamp_polish_support_repairs=1,
            uamp_polish_support_growth_steps=64,
            uamp_polish_support_shrink_steps=64,
            uamp_polish_support_shrink_batch_size=8,
            uamp_polish_support_prior=True,
            uamp_polish_support_prior_strength=1.0,
            uamp_polish_high_precision=True,
            uamp_polish_coordinate_steps=64,
            uamp_polish_coordinate_tolerance=1e-3,
            uamp_polish_noise_coordinate_steps=16,
            uamp_polish_noise_tolerance=1e-4,
            uamp_polish_noise_objective_tolerance=1e-10,
            uamp_polish_joint_cycles=8,
            uamp_polish_familywise_error_rate=0.05)
        settings.update(overrides)
        return cls.fast_hybrid_uamp(**settings)

    # --------------------------------------------------------- validation
    def validate(self) -> None:
        """Raise ValueError for invalid settings (solver.py:352-717)."""
        for rule, message in _RULES:
            if not rule(self):
                raise ValueError(message)


def _in(name, values):
    return lambda c: getattr(c, name) in values


_RULES = [
    (lambda c: c.max_iter >= 1 and c.min_iter >= 1,
     "Iteration counts must be positive."),
    (lambda c: 0.0 <= c.damping < 1.0, "damping must satisfy 0 <= damping < 1."),
    (_in("mode", ("auto", "dense", "cg", "uamp")),
     "mode must be 'auto', 'dense', 'cg', or 'uamp'."),
    (_in("noise_update", ("em", "evidence")),
     "noise_update must be 'em' or 'evidence'."),
    (_in("initialization", ("uniform", "backprojection")),
     "initialization must be 'uniform' or 'backprojection'."),
    (lambda c: 0.0 <= c.initialization_blend <= 1.0,
     "initialization_blend must be between 0 and 1."),
    (lambda c: c.fixed_noise_variance is None or c.fixed_noise_variance > 0,
     "fixed_noise_variance must be positive."),
    (lambda c: c.hutchinson_probes >= 1, "hutchinson_probes must be positive."),
    (lambda c: c.initial_hutchinson_probes >= 1,
     "initial_hutchinson_probes must be positive."),
    (lambda c: c.max_hutchinson_probes >= c.hutchinson_probes,
     "max_hutchinson_probes must be >= hutchinson_probes."),
    (lambda c: c.cg_tolerance > 0.0 and c.cg_initial_tolerance > 0.0,
     "CG tolerances must be positive."),
    (lambda c: c.cg_check_interval >= 1 and c.convergence_check_interval >= 1,
     "convergence check intervals must be positive."),
    (lambda c: c.cg_fixed_iterations >= 0,
     "cg_fixed_iterations cannot be negative."),
    (lambda c: c.cg_fixed_iterations <= c.cg_max_iter,
     "cg_fixed_iterations cannot exceed cg_max_iter."),
    (lambda c: c.cg_min_iter >= 1 and c.convergence_patience >= 1,
     "CG minimum iterations and patience must be positive."),
    (lambda c: c.cuda_cholesky_jitter_multiplier > 0.0,
     "cuda_cholesky_jitter_multiplier must be positive."),
    (_in("cg_matvec", ("tbt", "fourier")), "cg_matvec must be 'tbt' or 'fourier'."),
    (_in("cg_preconditioner", ("circulant", "low_rank", "nystrom", "none")),
     "cg_preconditioner must be 'circulant', 'low_rank', 'nystrom', or 'none'."),
    (lambda c: c.cg_preconditioner_rank >= 0,
     "cg_preconditioner_rank cannot be negative."),
    (lambda c: 0.0 <= c.cg_damping < 1.0,
     "cg_damping must satisfy 0 <= cg_damping < 1."),
    (lambda c: 0.0 <= c.uamp_damping < 1.0,
     "uamp_damping must satisfy 0 <= uamp_damping < 1."),
    (lambda c: c.uamp_damping <= c.uamp_max_damping < 1.0,
     "uamp_max_damping must be in [uamp_damping, 1)."),
    (lambda c: c.uamp_min_iter >= 1, "uamp_min_iter must be positive."),
    (lambda c: c.uamp_iteration_offset >= 0.0,
     "uamp_iteration_offset cannot be negative."),
    (lambda c: c.uamp_iteration_scale > 0.0,
     "uamp_iteration_scale must be positive."),
    (lambda c: c.uamp_polish_size >= 0, "uamp_polish_size cannot be negative."),
    (lambda c: c.uamp_polish_iterations >= 1,
     "uamp_polish_iterations must be positive."),
    (lambda c: 0 <= c.uamp_polish_residual_exchange_size <= 32,
     "uamp_polish_residual_exchange_size must be in [0, 32]."),
    (lambda c: c.uamp_polish_residual_exchange_size == 0 or (
        c.uamp_polish_iterations >= 2 and c.uamp_polish_prune_tail
        and c.uamp_polish_evidence_gate),
     "Residual support exchange requires at least two polish evaluations, "
     "uamp_polish_prune_tail=True and uamp_polish_evidence_gate=True."),
    (lambda c: 0.0 <= c.uamp_polish_damping < 1.0,
     "uamp_polish_damping must satisfy 0 <= damping < 1."),
    (lambda c: c.uamp_polish_evidence_margin_factor >= 0.0,
     "uamp_polish_evidence_margin_factor cannot be negative."),
    (lambda c: c.uamp_polish_max_truncation_delta is None
     or c.uamp_polish_max_truncation_delta > 0.0,
     "uamp_polish_max_truncation_delta must be positive."),
    (lambda c: 0 <= c.uamp_polish_support_repairs <= 4,
     "uamp_polish_support_repairs must be between 0 and 4."),
    (lambda c: c.uamp_polish_support_repairs == 0
     or c.uamp_polish_global_certificate,
     "uamp_polish_support_repairs requires uamp_polish_global_certificate=True."),
    (lambda c: 0 <= c.uamp_polish_support_growth_steps <= 64,
     "uamp_polish_support_growth_steps must be between 0 and 64."),
    (lambda c: c.uamp_polish_support_growth_steps == 0
     or c.uamp_polish_global_certificate,
     "uamp_polish_support_growth_steps requires "
     "uamp_polish_global_certificate=True."),
    (lambda c: 0 <= c.uamp_polish_support_shrink_steps <= 64,
     "uamp_polish_support_shrink_steps must be between 0 and 64."),
    (lambda c: c.uamp_polish_support_shrink_steps == 0
     or c.uamp_polish_global_certificate,
     "uamp_polish_support_shrink_steps requires "
     "uamp_polish_global_certificate=True."),
    (lambda c: 1 <= c.uamp_polish_support_shrink_batch_size <= 32,
     "uamp_polish_support_shrink_batch_size must be in [1, 32]."),
    (lambda c: c.uamp_polish_support_prior_strength >= 0.0,
     "uamp_polish_support_prior_strength cannot be negative."),
    (lambda c: not c.uamp_polish_support_prior
     or c.uamp_polish_global_certificate,
     "uamp_polish_support_prior requires uamp_polish_global_certificate=True."),
    (lambda c: 0 <= c.uamp_polish_coordinate_steps <= 64,
     "uamp_polish_coordinate_steps must be between 0 and 64."),
    (lambda c: c.uamp_polish_coordinate_tolerance >= 0.0,
     "uamp_polish_coordinate_tolerance cannot be negative."),
    (lambda c: c.uamp_polish_coordinate_steps == 0
     or c.uamp_polish_global_certificate,
     "uamp_polish_coordinate_steps requires "
     "uamp_polish_global_certificate=True."),
    (lambda c: 0 <= c.uamp_polish_noise_coordinate_steps <= 64,
     "uamp_polish_noise_coordinate_steps must be between 0 and 64."),
    (lambda c: c.uamp_polish_noise_tolerance > 0.0,
     "uamp_polish_noise_tolerance must be positive."),
    (lambda c: c.uamp_polish_noise_objective_tolerance >= 0.0,
     "uamp_polish_noise_objective_tolerance cannot be negative."),
    (lambda c: 1 <= c.uamp_polish_joint_cycles <= 32,
     "uamp_polish_joint_cycles must be between 1 and 32."),
    (lambda c: c.uamp_polish_noise_coordinate_steps == 0
     or c.uamp_polish_global_certificate,
     "uamp_polish_noise_coordinate_steps requires "
     "uamp_polish_global_certificate=True."),
    (lambda c: c.uamp_polish_noise_coordinate_steps == 0
     or c.fixed_noise_variance is None,
     "uamp_polish_noise_coordinate_steps cannot be used with "
     "fixed_noise_variance; either estimate noise or fix it."),
    (lambda c: c.uamp_polish_familywise_error_rate is None
     or 0.0 < c.uamp_polish_familywise_error_rate < 1.0,
     "uamp_polish_familywise_error_rate must be in (0, 1)."),
    (lambda c: c.uamp_polish_familywise_error_rate is None
     or c.uamp_polish_global_certificate,
     "uamp_polish_familywise_error_rate requires "
     "uamp_polish_global_certificate=True."),
    (lambda c: not c.uamp_polish_global_certificate
     or c.uamp_polish_prune_tail,
     "uamp_polish_global_certificate requires uamp_polish_prune_tail=True so "
     "inactive variances define the certified model."),
    (lambda c: not c.uamp_polish_fixed_budget_diagnostics or (
        c.uamp_polish_prune_tail and c.uamp_polish_evidence_gate),
     "uamp_polish_fixed_budget_diagnostics requires "
     "uamp_polish_prune_tail=True and uamp_polish_evidence_gate=True."),
    (lambda c: c.uamp_support_false_alarm_budget is None
     or c.uamp_support_false_alarm_budget > 0.0,
     "uamp_support_false_alarm_budget must be positive."),
    (_in("uamp_backend", ("auto", "torch", "triton")),
     "uamp_backend must be 'auto', 'torch', or 'triton'."),
    (lambda c: not (c.uamp_polish_high_precision
                    and c.uamp_backend == "triton"),
     "uamp_polish_high_precision=True requires the Torch backend: the "
     "fused UAMP kernels support complex64 only."),
    (_in("probe_type", ("rademacher", "hadamard")),
     "probe_type must be 'rademacher' or 'hadamard'."),
    (lambda c: c.active_set_start >= 1 and c.active_set_check_interval >= 1,
     "Active-set iteration counts must be positive."),
    (lambda c: c.active_set_min_size >= 1, "active_set_min_size must be positive."),
    (lambda c: c.active_set_max_size >= c.active_set_min_size,
     "active_set_max_size must be >= active_set_min_size."),
    (lambda c: 0.0 < c.active_set_tail_fraction < 1.0,
     "active_set_tail_fraction must be in (0, 1)."),
    (lambda c: 0.0 <= c.spatial_coupling <= 1.0,
     "spatial_coupling must be in [0, 1]."),
    (_in("spatial_boundary", ("zero", "circular")),
     "spatial_boundary must be 'zero' or 'circular'."),
    (lambda c: 0.0 < c.spatial_coupling_max <= 2.0,
     "spatial_coupling_max must be in (0, 2]."),
    (lambda c: 0.0 <= c.spatial_coupling_damping < 1.0,
     "spatial_coupling_damping must be in [0, 1)."),
    (lambda c: c.spatial_coupling_steps >= 1
     and c.spatial_coupling_update_interval >= 1,
     "Spatial coupling steps and update interval must be positive."),
    (lambda c: c.spatial_coupling_penalty >= 0.0
     and c.spatial_anisotropy_penalty >= 0.0,
     "Spatial coupling penalties cannot be negative."),
    (lambda c: 0.0 <= c.snapshot_covariance_damping < 1.0,
     "snapshot_covariance_damping must be in [0, 1)."),
    (lambda c: 0.0 <= c.snapshot_covariance_shrinkage < 1.0,
     "snapshot_covariance_shrinkage must be in [0, 1)."),
    (lambda c: 0.0 < c.snapshot_covariance_floor < 1.0,
     "snapshot_covariance_floor must be in (0, 1)."),
    (_in("snapshot_covariance_model", ("ar1", "full")),
     "snapshot_covariance_model must be 'ar1' or 'full'."),
    (lambda c: c.snapshot_covariance_start >= 0
     and c.snapshot_covariance_warmup >= 1,
     "Snapshot-covariance start must be nonnegative and warmup must be "
     "positive."),
    (lambda c: 0.0 <= c.snapshot_covariance_min_correlation < 1.0,
     "snapshot_covariance_min_correlation must be in [0, 1)."),
    (lambda c: c.snapshot_covariance_min_undersampling >= 1.0,
     "snapshot_covariance_min_undersampling must be at least one."),
    (_in("phase_error_axis", (0, 1)), "phase_error_axis must be 0 or 1."),
    (lambda c: c.phase_basis_rank >= 1 and c.phase_optimization_steps >= 1,
     "Phase basis rank and optimization steps must be positive."),
    (lambda c: c.phase_learning_rate > 0.0 and c.phase_smoothness >= 0.0,
     "Phase optimization settings are invalid."),
    (lambda c: c.phase_evidence_penalty >= 0.0,
     "phase_evidence_penalty cannot be negative."),
    (lambda c: c.phase_min_evidence_gain_per_parameter >= 0.0,
     "phase_min_evidence_gain_per_parameter cannot be negative."),
    (lambda c: c.phase_slq_probes >= 1 and c.phase_slq_steps >= 2,
     "SLQ requires positive probes and at least 2 steps."),
    (lambda c: c.phase_range_basis_rank >= 1,
     "phase_range_basis_rank must be positive."),
    (lambda c: c.migration_basis_rank >= 1,
     "migration_basis_rank must be positive."),
    (lambda c: c.migration_max_shift > 0.0 and c.migration_smoothness >= 0.0,
     "Range-migration settings are invalid."),
    (lambda c: c.motion_phase_amplitude_penalty >= 0.0
     and c.motion_migration_amplitude_penalty >= 0.0,
     "Motion-amplitude penalties cannot be negative."),
    (lambda c: c.motion_refinement_steps >= 0
     and c.motion_refinement_learning_rate > 0.0,
     "Motion-refinement settings are invalid."),
    (lambda c: c.offgrid_components is None or c.offgrid_components >= 1,
     "offgrid_components must be positive."),
    (lambda c: c.offgrid_max_components >= 1,
     "offgrid_max_components must be positive."),
    (lambda c: 0.0 < c.offgrid_relative_threshold < 1.0,
     "offgrid_relative_threshold must be in (0, 1)."),
    (lambda c: c.offgrid_minimum_separation >= 0.0,
     "offgrid_minimum_separation cannot be negative."),
    (lambda c: c.offgrid_max_iter >= 1 and c.offgrid_tolerance > 0.0
     and c.offgrid_max_step > 0.0 and c.offgrid_damping > 0.0
     and c.offgrid_ridge >= 0.0,
     "Off-grid refinement settings are invalid."),
]


@dataclass(slots=True)
class SBLResult:
    """Reconstruction and final posterior state (solver.py:720-792)."""

    x: Tensor
    posterior_variance: Tensor
    precision: Tensor
    noise_variance: Tensor
    effective_dof: Tensor
    iterations: int
    converged: bool
    relative_change: float
    mode: str
    arithmetic_dtype: str = ""
    uamp_backend_used: str | None = None
    active_set_size: int = 0
    offgrid: Any = None
    snapshot_covariance: Tensor | None = None
    snapshot_covariance_active: bool | None = None
    snapshot_snr_db: float | None = None
    learned_spatial_coupling: Tensor | None = None
    phase_error: Tensor | None = None
    phase_candidate: Tensor | None = None
    phase_correction_accepted: bool | None = None
    phase_evidence_gain: float | None = None
    range_migration: Tensor | None = None
    range_migration_candidate: Tensor | None = None
    negative_log_evidence: Tensor | None = None
    evidence_trace: Tensor | None = None
    truncation_delta_bound: Tensor | None = None
    hard_pruning_accepted: Tensor | None = None
    truncation_certified: Tensor | None = None
    inactive_max_evidence_ratio: Tensor | None = None
    inactive_best_index: Tensor | None = None
    inactive_evidence_gain: Tensor | None = None
    inactive_optimal_variance: Tensor | None = None
    support_repairs_accepted: Tensor | None = None
    support_exchanges_accepted: Tensor | None = None
    support_growth_accepted: Tensor | None = None
    support_shrink_accepted: Tensor | None = None
    active_max_evidence_gain: Tensor | None = None
    active_best_index: Tensor | None = None
    active_optimal_variance: Tensor | None = None
    active_max_deletion_gain: Tensor | None = None
    coordinate_steps_accepted: Tensor | None = None
    coordinate_stationary: Tensor | None = None
    inactive_ratio_threshold: Tensor | None = None
    familywise_support_certified: Tensor | None = None
    type2_coordinate_stationary: Tensor | None = None
    stationarity_ratio_threshold: Tensor | None = None
    familywise_ratio_threshold: Tensor | None = None
    familywise_no_discovery: Tensor | None = None
    active_model_selected: Tensor | None = None
    support_log_odds_penalty: Tensor | None = None
    penalized_negative_log_evidence: Tensor | None = None
    penalized_evidence_trace: Tensor | None = None
    inactive_penalized_evidence_gain: Tensor | None = None
    active_penalized_evidence_gain: Tensor | None = None
    type2_map_coordinate_stationary: Tensor | None = None
    noise_coordinate_steps_accepted: Tensor | None = None
    noise_evidence_gradient: Tensor | None = None
    noise_fixed_point_residual: Tensor | None = None
    noise_evidence_gain: Tensor | None = None
    noise_coordinate_stationary: Tensor | None = None
    joint_map_coordinate_stationary: Tensor | None = None
    joint_coordinate_cycles: Tensor | None = None
    fixed_budget_gamma_residual: Tensor | None = None
    fixed_budget_noise_residual: Tensor | None = None
    fixed_budget_objective_decrease: Tensor | None = None
    fixed_budget_monotone: Tensor | None = None
    support_mask: Tensor | None = None
    support_score_threshold: Tensor | None = None
    history: dict = field(default_factory=dict)


CONFIG_FIELDS = tuple(f.name for f in fields(SBLConfig))
