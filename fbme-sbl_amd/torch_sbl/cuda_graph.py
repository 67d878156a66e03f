"""Fixed-shape graph plans on MI355X (interface of torch_sbl/cuda_graph.py).

A plan builds the native FBME plan for the example's shape, runs ``warmup``
solves on a side stream, then captures one solve into a HIP graph through
``torch.cuda.CUDAGraph`` (hipGraph on ROCm).  ``replay`` copies new data into
the plan-owned input buffer and launches the graph; outputs are plan-owned and
overwritten by the next replay (torch_sbl/cuda_graph.py:12-20).
"""

from __future__ import annotations

import ctypes
from copy import deepcopy
from math import prod
from typing import Any, Sequence

import torch
from torch import Tensor

from . import _native
from .config import SBLConfig, SBLResult


class CUDAGraphSBLPlan:
    """Capture a fixed-shape SBL solve for low-latency replay
    (torch_sbl/cuda_graph.py:12-124)."""

    def __init__(self, example_data: Tensor, grid_shape: Sequence[int], *,
                 config: SBLConfig, warmup: int = 3) -> None:
        if not torch.cuda.is_available():
            raise RuntimeError("Graph execution requires a GPU.")
        if not isinstance(example_data, Tensor) or not example_data.is_cuda:
            raise ValueError("example_data must be a GPU tensor.")
        if warmup < 1:
            raise ValueError("warmup must be positive.")
        if config.return_history or config.verbose:
            raise ValueError(
                "Graph plans require return_history=False and verbose=False.")
        if config.tol > 0.0 or config.eager_change_tracking:
            raise ValueError(
                "Graph plans require a fixed iteration budget: set tol=0 and "
                "eager_change_tracking=False.")
        if (config.phase_error_correction or config.range_migration_correction
                or config.offgrid_refinement or config.learn_snapshot_covariance):
            raise ValueError(
                "Capture the core SBL stage only; motion, off-grid, and "
                "snapshot-covariance refinement are not graph-static.")
        config.validate()
        if example_data.dtype != torch.complex64:
            raise NotImplementedError("Graph plans run complex64 data.")
        self.grid_shape = tuple(int(v) for v in grid_shape)
        if len(self.grid_shape) != 2:
            raise ValueError("grid_shape must be (N1, N2).")
        self.config = deepcopy(config)
        self.device = example_data.device
        self.dtype = example_data.dtype
        self.shape = tuple(example_data.shape)
        if len(self.shape) < 2:
            raise ValueError("example_data must end with (M1, M2).")
        samples = self.shape[-2:]
        if any(m > n for m, n in zip(samples, self.grid_shape)):
            raise ValueError(
                "Observed dimensions cannot exceed reconstruction dimensions.")
        batch_shape = self.shape[:-2]
        batch = prod(batch_shape) if batch_shape else 1
        # mode resolution of fast_sbl (solver.py:1025-1036)
        mode = self.config.mode
        if mode == "auto":
            mode = ("dense" if prod(samples) <= self.config.dense_threshold
                    else "cg")
        self._solver = None
        if mode in ("cg", "dense"):
            self._capture_general(example_data, samples, batch, batch_shape,
                                  mode == "dense", warmup)
            return
        _native.check_supported(config)
        opt = _native.plan_options(self.config, self.grid_shape, samples)
        shapes = [(self.grid_shape[0], self.grid_shape[1], samples[0],
                   samples[1])] * batch
        opt = _native.tuned(shapes, opt, self.device)
        with torch.cuda.device(self.device):
            self._plan = _native.NativePlan(shapes, opt, self.device)
            self._views = _native.ResultViews(self._plan, self.config,
                                              batch_shape, self.grid_shape)
            self.input_buffer = torch.empty_like(example_data)
            self.input_buffer.copy_(example_data)
            flat = self.input_buffer.view(-1)
            side = torch.cuda.Stream(device=self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                for _ in range(warmup):
                    self._plan.run(flat)
                    self._views.refresh()
            torch.cuda.current_stream(self.device).wait_stream(side)
            torch.cuda.synchronize(self.device)
            self._graph = torch.cuda.CUDAGraph()
            try:
                with torch.cuda.graph(self._graph):
                    self._plan.run(flat)
                    self._views.refresh()
            except Exception as error:
                raise RuntimeError(
                    "Graph capture failed for this SBL configuration.") from error
        self._result = self._views.result()

    def _capture_general(self, example_data, samples, batch, batch_shape,
                         dense, warmup):
        """mode='cg' with a fixed PCG budget, or mode='dense': the general
        solver's EM loop captured once (torch_sbl.pcg.solve).  Solves whose
        loop reads the device (PCG early exit, active-set switch,
        convergence checks) cannot be captured, as in the reference, where
        the capture itself fails."""
        from . import pcg
        why = pcg.capturable(self.config, batch, dense)
        if why is not None:
            raise RuntimeError(
                "Graph capture failed for this SBL configuration: " + why + ".")
        pcg._check_supported(self.config)
        with torch.cuda.device(self.device):
            self.input_buffer = torch.empty_like(example_data)
            self.input_buffer.copy_(example_data)
            flat = self.input_buffer.view(batch, *samples)
            side = torch.cuda.Stream(device=self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                self._solver = pcg._Solver(flat, self.grid_shape, self.config,
                                           dense=dense)
                for _ in range(warmup):
                    pcg.solve(self._solver, batch_shape, capture=True)
            torch.cuda.current_stream(self.device).wait_stream(side)
            torch.cuda.synchronize(self.device)
            self._graph = torch.cuda.CUDAGraph()
            try:
                with torch.cuda.graph(self._graph):
                    self._solver.stream = ctypes.c_void_p(
                        torch.cuda.current_stream(self.device).cuda_stream)
                    self._result = pcg.solve(self._solver, batch_shape,
                                             capture=True)
            except Exception as error:
                raise RuntimeError(
                    "Graph capture failed for this SBL configuration.") from error

    @property
    def result(self) -> SBLResult:
        """Plan-owned output of the most recent replay."""
        return self._result

    @property
    def native_plan(self):
        if self._solver is not None:
            raise AttributeError("general-solver graph plans have no native FBME plan")
        return self._plan

    def factorization_failed(self) -> bool:
        """Dense mode: whether the last replay's FP64 Cholesky met a
        non-positive pivot after its loading (the condition solver.py:2971
        raises on; a replay cannot raise, so it is read here)."""
        if self._solver is None or not self._solver.dense:
            return False
        return bool(self._solver.status.any())

    def replay(self, data: Tensor | None = None, *,
               synchronize: bool = False) -> SBLResult:
        """Replay the captured solve, optionally copying new GPU input."""
        if data is not None:
            self._validate_input(data)
            self.input_buffer.copy_(data)
        self._graph.replay()
        if synchronize:
            torch.cuda.synchronize(self.device)
        return self._result

    def _validate_input(self, data: Tensor) -> None:
        if not isinstance(data, Tensor) or not data.is_cuda:
            raise ValueError("Replay data must be a GPU tensor.")
        if data.device != self.device:
            raise ValueError(
                f"Replay data must be on {self.device}, got {data.device}.")
        if data.dtype != self.dtype:
            raise ValueError(
                f"Replay data must have dtype {self.dtype}, got {data.dtype}.")
        if tuple(data.shape) != self.shape:
            raise ValueError(
                f"Replay data must have shape {self.shape}, got "
                f"{tuple(data.shape)}.")


class RealtimeSBLPlan(CUDAGraphSBLPlan):
    """Strict fixed-budget real-time plan (torch_sbl/cuda_graph.py:127-216)."""

    profile = "fbme-sbl-fixed-budget"
    certificate_scope = "fixed-budget-monotone-type2"

    def __init__(self, example_data: Tensor, grid_shape: Sequence[int], *,
                 polish_size: int = 32, warmup: int = 3,
                 config_overrides: dict[str, Any] | None = None) -> None:
        overrides: dict[str, Any] = {
            "device": example_data.device,
            "uamp_polish_size": polish_size,
        }
        if config_overrides is not None:
            overrides.update(config_overrides)
        config = SBLConfig.realtime_hybrid_uamp(**overrides)
        self._validate_realtime_config(config)
        super().__init__(example_data, grid_shape, config=config,
                         warmup=warmup)

    @staticmethod
    def _validate_realtime_config(config: SBLConfig) -> None:
        if not config.uamp_polish_fixed_budget_diagnostics:
            raise ValueError(
                "RealtimeSBLPlan requires fixed-budget diagnostics so the "
                "paper objective and residuals are always reported.")
        if any(v != 0.0 for v in (config.hyper_shape, config.hyper_rate,
                                  config.noise_shape, config.noise_rate)):
            raise ValueError(
                "RealtimeSBLPlan requires zero numerical hyperpriors so its "
                "exact evidence gate matches the stated paper objective.")
        forbidden = {
            "uamp_polish_global_certificate":
                config.uamp_polish_global_certificate,
            "uamp_polish_support_repairs":
                config.uamp_polish_support_repairs != 0,
            "uamp_polish_support_growth_steps":
                config.uamp_polish_support_growth_steps != 0,
            "uamp_polish_support_shrink_steps":
                config.uamp_polish_support_shrink_steps != 0,
            "uamp_polish_coordinate_steps":
                config.uamp_polish_coordinate_steps != 0,
            "uamp_polish_noise_coordinate_steps":
                config.uamp_polish_noise_coordinate_steps != 0,
            "uamp_polish_high_precision": config.uamp_polish_high_precision,
        }
        enabled = [name for name, value in forbidden.items() if value]
        if enabled:
            raise ValueError(
                "RealtimeSBLPlan excludes dynamic research certificates; "
                "disable: " + ", ".join(enabled))
        if config.uamp_backend != "triton":
            raise ValueError(
                "RealtimeSBLPlan requires the fused-kernel UAMP backend "
                "(uamp_backend='triton' selects it here too).")


def make_cuda_graph_sbl_plan(example_data: Any, grid_shape: Sequence[int], *,
                             config: SBLConfig | None = None,
                             warmup: int = 3) -> CUDAGraphSBLPlan:
    """Convenience constructor (torch_sbl/cuda_graph.py:219-236)."""
    tensor = torch.as_tensor(example_data)
    active = config or SBLConfig.fast_hybrid_uamp(device=tensor.device)
    return CUDAGraphSBLPlan(tensor, grid_shape, config=active, warmup=warmup)


def make_realtime_sbl_plan(example_data: Any, grid_shape: Sequence[int], *,
                           polish_size: int = 32, warmup: int = 3,
                           config_overrides: dict[str, Any] | None = None
                           ) -> RealtimeSBLPlan:
    """Strict fixed-budget real-time plan (torch_sbl/cuda_graph.py:239-255)."""
    tensor = torch.as_tensor(example_data)
    return RealtimeSBLPlan(tensor, grid_shape, polish_size=polish_size,
                           warmup=warmup, config_overrides=config_overrides)
