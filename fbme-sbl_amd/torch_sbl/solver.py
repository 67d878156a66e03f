"""Eager entry points fast_sbl / sbl_2d / sbl_1d on the MI355X kernels.

Interface of torch_sbl/solver.py:813-931.  Leading dimensions of ``data`` are
independent scenes solved together.  Execution modes:

* ``mode="uamp"`` runs the native FBME plan (HIP kernels, one C call);
* ``mode="cg"`` runs the covariance-free PCG + Hutchinson path on HIP
  kernels (torch_sbl.pcg);
* ``mode="dense"`` (and ``"auto"`` resolving to it when M <= 2048) is the
  exact observation-domain EM (FP64 tiled Cholesky of the M x M covariance on
  HIP kernels; M <= 2048, grid axes <= 1024; with joint sparsity one FP64
  posterior per snapshot).

``mode="uamp"`` with ``tol > 0`` (or history / change tracking) drives the
plan's UAMP iterations from the host and stops at the reference's
convergence checks; ``uamp_polish_size=0`` returns the UAMP posterior
without a polish.

Only GPU tensors are accepted: there is no CPU execution path.
"""

from __future__ import annotations

from math import prod
from typing import Any, Sequence

import torch

from . import _native
from .config import SBLConfig, SBLResult

__all__ = ["SBLConfig", "SBLResult", "fast_sbl", "sbl_1d", "sbl_2d",
           "clear_operator_cache"]


def clear_operator_cache() -> None:
    """Release cached plans (torch_sbl/solver.py:842-844)."""
    _native.clear_cache()
    try:
        from . import pcg
        pcg.clear_cache()
    except ImportError:
        pass


def sbl_1d(data: Any, grid_size: int, *, config: SBLConfig | None = None,
           observation_mask: Any | None = None) -> SBLResult:
    """1-D spectra: data [..., M] -> x [..., grid_size] (solver.py:847-867).
    Runs as a 1 x N grid."""
    return fast_sbl(data, (int(grid_size),), config=config,
                    observation_mask=observation_mask)


def _squeeze_1d(res: SBLResult) -> SBLResult:
    """A 1-D grid ran as 1 x N: grid-shaped fields lose the unit axis, so
    they are [..., N] as the reference returns them."""
    for name in ("x", "posterior_variance", "precision", "support_mask"):
        value = getattr(res, name, None)
        if value is not None:
            setattr(res, name, value.squeeze(-2))
    return res


def sbl_2d(data: Any, grid_shape: Sequence[int], *,
           config: SBLConfig | None = None,
           observation_mask: Any | None = None) -> SBLResult:
    """2-D scenes: data [..., M1, M2] -> x [..., N1, N2]
    (solver.py:885-922)."""
    shape = tuple(int(v) for v in grid_shape)
    if len(shape) != 2:
        raise ValueError("sbl_2d requires grid_shape=(N1, N2).")
    if config is not None and (config.phase_error_correction
                               or config.learn_snapshot_covariance):
        raise NotImplementedError(
            "autofocus and correlated-MMV reconstruction are outside the "
            "FBME hot path on MI355X.")
    return fast_sbl(data, shape, config=config,
                    observation_mask=observation_mask)


def _resolve_device(requested, tensor: torch.Tensor) -> torch.device:
    if str(requested) != "auto":
        device = torch.device(requested)
        if device.type == "cuda" and device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        return device
    if tensor.device.type != "cpu":
        return tensor.device
    if torch.cuda.is_available():
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device("cpu")


def fast_sbl(data: Any, grid_shape: Sequence[int], *,
             config: SBLConfig | None = None,
             observation_mask: Any | None = None) -> SBLResult:
    """Fourier SBL without materialising the dictionary
    (torch_sbl/solver.py:925-2393)."""
    cfg = config or SBLConfig()
    cfg.validate()
    if cfg.learn_snapshot_covariance:
        raise ValueError(
            "learn_snapshot_covariance is available through sbl_2d only.")
    if cfg.phase_error_correction:
        raise ValueError(
            "phase_error_correction is available through sbl_2d only.")
    grid = tuple(int(v) for v in grid_shape)
    if len(grid) not in (1, 2):
        raise ValueError("grid_shape must contain one or two dimensions.")
    one_d = len(grid) == 1
    if one_d:
        grid = (1, grid[0])
        data = torch.as_tensor(data).unsqueeze(-2)
        if observation_mask is not None:
            observation_mask = torch.as_tensor(observation_mask).unsqueeze(0)
    if (cfg.uamp_support_false_alarm_budget is not None
            and cfg.uamp_support_false_alarm_budget >= prod(grid)):
        raise ValueError(
            "uamp_support_false_alarm_budget must be smaller than the "
            "reconstruction grid size.")
    tensor = torch.as_tensor(data)
    if tensor.ndim < 2:
        raise ValueError("data has fewer dimensions than grid_shape.")
    samples = tuple(int(v) for v in tensor.shape[-2:])
    batch_shape = tuple(int(v) for v in tensor.shape[:-2])
    if any(m > n for m, n in zip(samples, grid)):
        raise ValueError(
            "Observed dimensions cannot exceed reconstruction dimensions.")
    device = _resolve_device(cfg.device, tensor)
    if device.type != "cuda":
        raise RuntimeError(
            "FBME runs on MI355X GPUs only here; pass a GPU tensor or "
            "device (the CPU oracle lives in the test suite).")
    mask = None
    if observation_mask is not None:
        # solver.py:986-1000
        mask = torch.as_tensor(observation_mask, dtype=torch.bool)
        if tuple(mask.shape) != samples:
            raise ValueError(
                "observation_mask must match the trailing data dimensions.")
        if not bool(mask.any()):
            raise ValueError(
                "observation_mask must retain at least one sample.")
        if cfg.offgrid_refinement:
            raise ValueError(
                "Run off-grid refinement after masked reconstruction.")
    if tensor.dtype in (torch.complex128, torch.float64):
        raise NotImplementedError(
            "complex128 arithmetic is not provided by the MI355X kernels; "
            "pass complex64 data.")
    y = tensor.to(device=device, dtype=torch.complex64)
    batch = prod(batch_shape) if batch_shape else 1
    m = prod(samples) if mask is None else int(mask.sum())
    threshold = cfg.dense_threshold if device.type == "cuda" else min(
        cfg.dense_threshold, 512)
    mode = ("dense" if cfg.mode == "auto" and m <= threshold else
            "cg" if cfg.mode == "auto" else cfg.mode)
    if mode in ("cg", "dense"):
        from . import pcg
        res = pcg.sbl_cg(y.reshape(batch, *samples), grid, cfg,
                         batch_shape=batch_shape, observation_mask=mask,
                         dense=mode == "dense")
        return _finish(res, y, grid, cfg, one_d)
    if cfg.uamp_backend == "torch":
        raise NotImplementedError(
            "uamp_backend='torch' selects the reference's PyTorch UAMP; this "
            "build always runs the HIP kernels.")
    _native.check_supported(cfg)
    opt = _native.plan_options(cfg, grid, samples, None if mask is None else m)
    shapes = [(grid[0], grid[1], samples[0], samples[1])] * batch
    plan = _native.cached_plan(shapes, opt, device, mask)
    views = _native.ResultViews(plan, cfg, batch_shape, grid)
    with torch.cuda.device(device):
        flat = y.reshape(-1).contiguous()
        if (cfg.tol > 0.0 or cfg.return_history or cfg.eager_change_tracking
                or cfg.verbose):
            # convergence checks / history: the loop is driven from the host
            _native.run_eager(plan, views, flat, cfg)
        else:
            plan.run(flat)
            views.refresh()
        # Eager calls return fresh tensors; graph plans return plan storage.
        return _finish(views.result(clone=True), y, grid, cfg, one_d)


def _finish(res: SBLResult, y: torch.Tensor, grid, cfg: SBLConfig, one_d: bool) -> SBLResult:
    """Result shapes of the caller's grid, then the off-grid refinement of
    the reconstruction on that grid (solver.py:2139-2153)."""
    if one_d:
        res = _squeeze_1d(res)
        y = y.squeeze(-2)
        grid = (grid[1],)
    if not cfg.offgrid_refinement:
        return res
    from .offgrid import refine_offgrid_fourier
    res.offgrid = refine_offgrid_fourier(
        y, res.x, grid, component_count=cfg.offgrid_components,
        max_components=cfg.offgrid_max_components,
        relative_threshold=cfg.offgrid_relative_threshold,
        minimum_separation=cfg.offgrid_minimum_separation,
        max_iter=cfg.offgrid_max_iter, tolerance=cfg.offgrid_tolerance,
        max_step=cfg.offgrid_max_step, damping=cfg.offgrid_damping,
        ridge=cfg.offgrid_ridge)
    return res
