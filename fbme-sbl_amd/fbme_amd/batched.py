"""Uneven batches: scenes of different grid and sampling shapes in one plan.

One native plan holds every scene back to back in the workspace; each pass
of the solve is one launch over all scenes (per-scene offsets, radix chains
and UAMP budgets in the scene table; scenes whose budget T_u is spent drop
out of the later sweeps, which is the per-scene convergence mask), and the
whole solve is captured in one hipGraph.  The pass shapes come from the
autotuner's table for this workload and GPU (``tune=True`` runs the tuner
first).  ``masks`` optionally gives scenes observation masks [M1, M2]
(None for a fully observed scene).  Outputs are plan-owned and overwritten by the next replay, as in
the reference's graph plans (torch_sbl/cuda_graph.py:17-19).
"""

from __future__ import annotations

from typing import Sequence

import torch

from . import _lib, autotune
from .plan import NativePlan


class UnevenPlan:
    """Graph-captured solve of a fixed list of (N1, N2, M1, M2) scenes."""

    def __init__(self, shapes: Sequence[Sequence[int]], *, device, profile: str = "fast", warmup: int = 2,
                 tune: bool = False, options: _lib.FbmePlanOptions | None = None,
                 masks: Sequence | None = None) -> None:
        if profile not in ("fast", "stable"):
            raise ValueError("profile must be 'fast' or 'stable'")
        if warmup < 1:
            raise ValueError("warmup must be positive")
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("UnevenPlan needs a GPU device")
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        stable = profile == "stable"
        if tune:
            self.tuning = autotune.tune_plan(self.shapes, stable, self.device)
        opt = options if options is not None else _lib.default_options(stable)
        if options is None:
            opt = autotune.tuned_plan_options(self.shapes, stable, self.device, opt)
        self.options = opt
        with torch.cuda.device(self.device):
            self.plan = NativePlan(self.shapes, opt, self.device, masks)
            self.input = torch.zeros(self.plan.total_samples, dtype=torch.complex64, device=self.device)
            side = torch.cuda.Stream(device=self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                for _ in range(warmup):
                    self.plan.run(self.input)
            torch.cuda.current_stream(self.device).wait_stream(side)
            torch.cuda.synchronize(self.device)
            self.graph = torch.cuda.CUDAGraph()
            try:
                with torch.cuda.graph(self.graph):
                    self.plan.run(self.input)
            except Exception as error:
                raise RuntimeError("graph capture of the uneven plan failed") from error
        self._slices = self.plan.scene_slices()

    @property
    def nscenes(self) -> int:
        return len(self.shapes)

    def pack(self, data: Sequence[torch.Tensor]) -> None:
        """Copy per-scene observations [M1_i, M2_i] into the plan's input."""
        if len(data) != self.nscenes:
            raise ValueError(f"expected {self.nscenes} scenes, got {len(data)}")
        for i, (y, (n1, n2, m1, m2)) in enumerate(zip(data, self.shapes)):
            if not isinstance(y, torch.Tensor) or y.device != self.device or y.dtype != torch.complex64:
                raise ValueError(f"scene {i}: data must be complex64 on {self.device}")
            if tuple(y.shape) != (m1, m2):
                raise ValueError(f"scene {i}: expected shape {(m1, m2)}, got {tuple(y.shape)}")
            so = self._slices[i][1]
            self.input[so:so + m1 * m2].copy_(y.reshape(-1))

    def replay(self, data: Sequence[torch.Tensor] | torch.Tensor | None = None, *,
               synchronize: bool = False):
        """Replay the captured solve; `data` is a list of per-scene tensors or
        the packed 1-D input (scenes in caller order)."""
        if isinstance(data, torch.Tensor):
            if data.shape != self.input.shape or data.dtype != self.input.dtype or data.device != self.device:
                raise ValueError(f"packed data must be complex64 [{self.input.numel()}] on {self.device}")
            self.input.copy_(data)
        elif data is not None:
            self.pack(data)
        self.graph.replay()
        if synchronize:
            torch.cuda.synchronize(self.device)
        return self.plan.out

    def scene(self, i: int) -> dict:
        """Views of scene i's outputs (plan storage)."""
        n1, n2, _, _ = self.shapes[i]
        go = self._slices[i][0]
        o = self.plan.out
        n = n1 * n2
        return dict(x=o.x[go:go + n].view(n1, n2),
                    posterior_variance=o.posterior_variance[go:go + n].view(n1, n2),
                    precision=o.precision[go:go + n].view(n1, n2),
                    support_mask=o.support_mask[go:go + n].view(n1, n2),
                    noise_variance=o.noise[i], effective_dof=o.dof[i], status=o.status[i])
