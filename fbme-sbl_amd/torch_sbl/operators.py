"""Implicit prefix-DFT operator on MI355X (interface of
torch_sbl/operators.py:23-501 for the methods the FBME hot path uses).

``forward_flat`` / ``adjoint_flat`` and the ``[B, N, R]`` forms run the
pruned two-pass FFT kernels of libfbme_hip.so (fbme_prefix_forward /
fbme_prefix_adjoint); the DFT is unnormalised, A = fft2(.)[:M1, :M2] and
A^H = N ifft2(zero-pad(.)).  With an observation mask the operator keeps
the observed rows (operators.py:57-75, 473-496): forward gathers them
from the full prefix block, adjoint scatters them into zeros first.
"""

from __future__ import annotations

import ctypes
from math import prod
from typing import Sequence

import torch
from torch import Tensor

from fbme_amd import _lib


def _next_smooth_length(minimum: int) -> int:
    """First length >= minimum with only 2/3/5 factors (the FFT kernels'
    radices).  The reference also admits 7 (operators.py:10-20)."""
    candidate = int(minimum)
    while True:
        rem = candidate
        for f in (2, 3, 5):
            while rem % f == 0:
                rem //= f
        if rem == 1:
            return candidate
        candidate += 1


class PrefixFourierOperator:
    """Leading M1 x M2 block of the unnormalised 2-D DFT of an N1 x N2 grid."""

    def __init__(self, grid_shape: Sequence[int], sample_shape: Sequence[int],
                 *, device: torch.device, complex_dtype: torch.dtype = torch.complex64,
                 observation_mask: Tensor | None = None) -> None:
        grid = tuple(int(v) for v in grid_shape)
        samples = tuple(int(v) for v in sample_shape)
        if len(grid) == 1:
            grid, samples = (1, grid[0]), (1, samples[0])
        if len(grid) != 2 or len(samples) != 2:
            raise ValueError("Only 1-D and 2-D Fourier grids are supported.")
        if any(m <= 0 or m > n for m, n in zip(samples, grid)):
            raise ValueError("Each sample dimension must satisfy 0 < M <= N.")
        if complex_dtype != torch.complex64:
            raise NotImplementedError("the HIP operator is complex64 only.")
        self.grid_shape = grid
        self.sample_shape = samples
        self.device = torch.device(device)
        self.complex_dtype = complex_dtype
        self.real_dtype = torch.float32
        self.n_grid = prod(grid)
        self.n_samples = prod(samples)
        self.full_n_samples = self.n_samples
        self.observation_mask = None
        self.observed_indices = None
        if observation_mask is not None:
            mask = torch.as_tensor(observation_mask, dtype=torch.bool).to(self.device)
            if tuple(mask.shape) != samples:
                raise ValueError(f"observation_mask must have shape {samples}, got {tuple(mask.shape)}.")
            indices = mask.reshape(-1).nonzero(as_tuple=False).squeeze(1)
            if indices.numel() == 0:
                raise ValueError("observation_mask must retain at least one sample.")
            self.observation_mask = mask
            self.observed_indices = indices
            self.n_samples = int(indices.numel())
        self.embedding_shape = tuple(_next_smooth_length(2 * m - 1)
                                     for m in samples)
        self._mid: Tensor | None = None

    @property
    def lipschitz_upper_bound(self) -> float:
        """||A||_2^2 = N for a set of unnormalised DFT rows."""
        return float(self.n_grid)

    @property
    def is_complete(self) -> bool:
        return self.observed_indices is None

    def restrict(self, full_data: Tensor) -> Tensor:
        """[B, M1 M2, ...] -> [B, M_observed, ...] (operators.py:473-482)."""
        if self.observed_indices is None:
            return full_data
        return full_data.index_select(1, self.observed_indices)

    def expand(self, observed_data: Tensor) -> Tensor:
        """[B, M_observed, ...] -> [B, M1 M2, ...] with zeros on the
        unobserved samples (operators.py:484-496)."""
        if self.observed_indices is None:
            return observed_data
        full = torch.zeros((observed_data.shape[0], self.full_n_samples, *observed_data.shape[2:]),
                           dtype=observed_data.dtype, device=observed_data.device)
        full.index_copy_(1, self.observed_indices, observed_data)
        return full

    def _workspace(self, batch: int) -> Tensor:
        need = batch * self.grid_shape[0] * self.sample_shape[1]
        if self._mid is None or self._mid.numel() < need:
            self._mid = torch.empty(need, dtype=torch.complex64,
                                    device=self.device)
        return self._mid

    def _check(self, t: Tensor, length: int, name: str) -> None:
        if t.ndim != 2 or t.shape[1] != length:
            raise ValueError(f"{name} must have shape [B, {length}].")
        if t.dtype != torch.complex64 or t.device != self.device:
            raise ValueError(f"{name} must be complex64 on {self.device}.")

    def forward_flat(self, coefficients: Tensor) -> Tensor:
        """A x for x [B, N] -> [B, M] (operators.py:237-253)."""
        self._check(coefficients, self.n_grid, "coefficients")
        x = coefficients.contiguous()
        b = x.shape[0]
        y = torch.empty((b, self.full_n_samples), dtype=torch.complex64,
                        device=self.device)
        mid = self._workspace(b)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.load().fbme_prefix_forward(
            ctypes.c_void_p(stream), b, *self.grid_shape, *self.sample_shape,
            x.data_ptr(), y.data_ptr(), mid.data_ptr()), "fbme_prefix_forward")
        return self.restrict(y)

    def adjoint_flat(self, data: Tensor) -> Tensor:
        """A^H v for v [B, M] -> [B, N] (operators.py:255-280)."""
        self._check(data, self.n_samples, "data")
        v = self.expand(data).contiguous()
        b = v.shape[0]
        x = torch.empty((b, self.n_grid), dtype=torch.complex64,
                        device=self.device)
        mid = self._workspace(b)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.load().fbme_prefix_adjoint(
            ctypes.c_void_p(stream), b, *self.grid_shape, *self.sample_shape,
            v.data_ptr(), x.data_ptr(), mid.data_ptr()), "fbme_prefix_adjoint")
        return x

    def forward(self, coefficients: Tensor) -> Tensor:
        """[B, N, R] -> [B, M, R] (operators.py:208-216)."""
        if coefficients.ndim != 3 or coefficients.shape[1] != self.n_grid:
            raise ValueError(f"coefficients must have shape [B, {self.n_grid}, R].")
        b, _, r = coefficients.shape
        flat = coefficients.permute(0, 2, 1).reshape(b * r, self.n_grid)
        out = self.forward_flat(flat)
        return out.reshape(b, r, self.n_samples).permute(0, 2, 1)

    def adjoint(self, data: Tensor) -> Tensor:
        """[B, M, R] -> [B, N, R] (operators.py:218-235)."""
        if data.ndim != 3 or data.shape[1] != self.n_samples:
            raise ValueError(f"data must have shape [B, {self.n_samples}, R].")
        b, _, r = data.shape
        flat = data.permute(0, 2, 1).reshape(b * r, self.n_samples)
        out = self.adjoint_flat(flat)
        return out.reshape(b, r, self.n_grid).permute(0, 2, 1)

    def fourier_columns(self, flat_indices: Tensor) -> Tensor:
        """Selected DFT columns A[:, S] -> [B, M, K] with the phase reduced
        modulo N in integers (exact), unlike the float32 phase of
        operators.py:444-466."""
        if flat_indices.ndim != 2:
            raise ValueError("flat_indices must have shape [B, K].")
        n1, n2 = self.grid_shape
        m1, m2 = self.sample_shape
        k1 = torch.div(flat_indices, n2, rounding_mode="floor")
        k2 = torch.remainder(flat_indices, n2)
        a1 = torch.arange(m1, device=flat_indices.device)
        a2 = torch.arange(m2, device=flat_indices.device)
        r1 = torch.remainder(a1[None, :, None] * k1[:, None, :], n1)
        r2 = torch.remainder(a2[None, :, None] * k2[:, None, :], n2)
        phase = (r1[:, :, None, :].double() / n1
                 + r2[:, None, :, :].double() / n2)
        cols = torch.exp(-2j * torch.pi * phase).to(torch.complex64)
        return self.restrict(cols.reshape(flat_indices.shape[0], m1 * m2, -1))
