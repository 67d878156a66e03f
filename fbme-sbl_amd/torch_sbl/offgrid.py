"""Off-grid refinement of on-grid SBL peaks on MI355X (interface of
torch_sbl/offgrid.py:11-226).

``refine_offgrid_fourier`` runs one HIP workgroup per batch item
(``fbme_offgrid_refine`` in libfbme_hip.so): peak selection, the BIC model
order, projected Gauss--Newton and the position standard deviations all
happen in the kernel, in FP64, without forming the M x K dictionary.  Same
arguments, validation and result fields as the reference; results are
returned in the data's real / complex precision.  1-D grids run as 1 x N
(fbme_offgrid_refine_ws with one position coordinate per component); at
most 64 candidate and refined components per item (the default
max_components).  Items with more than 32 refined components keep their
refinement matrices in a device workspace of 388 KiB per item.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from math import prod
from typing import Sequence

import torch
from torch import Tensor

from fbme_amd import _lib

_KMAX = 64        # refined components (more than _KLDS: with a workspace)
_KLDS = 32
_CANDIDATES = 64


@dataclass(slots=True)
class OffGridResult:
    """Continuous Fourier parameters refined from an on-grid SBL estimate
    (offgrid.py:11-26).  Positions are in grid-bin units."""

    initial_positions: Tensor
    positions: Tensor
    amplitudes: Tensor
    position_std: Tensor
    initial_residual_norm: Tensor
    residual_norm: Tensor
    iterations: Tensor
    converged: Tensor


def continuous_fourier_dictionary(positions: Tensor, grid_shape: Sequence[int],
                                  sample_shape: Sequence[int]) -> Tensor:
    """Atoms exp(-2 pi i sum_d m_d p_d / N_d) at continuous positions [K, D]
    -> [M, K] (offgrid.py:29-67)."""
    grid_shape = tuple(int(v) for v in grid_shape)
    sample_shape = tuple(int(v) for v in sample_shape)
    if len(grid_shape) not in (1, 2):
        raise ValueError("Only 1-D and 2-D Fourier grids are supported.")
    if len(grid_shape) != len(sample_shape):
        raise ValueError("grid_shape and sample_shape must have equal rank.")
    if positions.ndim != 2 or positions.shape[1] != len(grid_shape):
        raise ValueError("positions must have shape [K, len(grid_shape)].")
    real = torch.float64 if positions.dtype == torch.float64 else torch.float32
    axes = [torch.arange(n, device=positions.device, dtype=real) for n in sample_shape]
    coords = torch.stack([g.reshape(-1) for g in torch.meshgrid(*axes, indexing="ij")], dim=1)
    scale = torch.tensor(grid_shape, device=positions.device, dtype=real)
    phase = (coords / scale) @ positions.to(real).mT
    cplx = torch.complex128 if real == torch.float64 else torch.complex64
    return torch.exp((-2j * torch.pi * phase).to(cplx))


def refine_offgrid_fourier(data: Tensor, coefficients: Tensor, grid_shape: Sequence[int], *,
                           component_count: int | None = None, max_components: int = 64,
                           relative_threshold: float = 0.1, minimum_separation: float = 1.5,
                           max_iter: int = 15, tolerance: float = 1e-4, max_step: float = 0.35,
                           damping: float = 1e-3, ridge: float = 1e-7) -> OffGridResult:
    """Refine on-grid SBL peaks by projected Gauss--Newton
    (offgrid.py:70-226)."""
    grid_shape = tuple(int(v) for v in grid_shape)
    if len(grid_shape) not in (1, 2):
        raise ValueError("grid_shape must be one- or two-dimensional.")
    if coefficients.ndim < len(grid_shape):
        raise ValueError("coefficients has fewer dimensions than grid_shape.")
    if tuple(coefficients.shape[-len(grid_shape):]) != grid_shape:
        raise ValueError("The trailing coefficient shape must equal grid_shape.")
    if data.ndim < len(grid_shape):
        raise ValueError("data has fewer dimensions than grid_shape.")
    sample_shape = tuple(int(v) for v in data.shape[-len(grid_shape):])
    if any(m > n for m, n in zip(sample_shape, grid_shape)):
        raise ValueError("Observed dimensions cannot exceed grid dimensions.")
    if component_count is not None and component_count < 1:
        raise ValueError("component_count must be positive.")
    if max_components < 1:
        raise ValueError("max_components must be positive.")
    if not 0.0 < relative_threshold < 1.0:
        raise ValueError("relative_threshold must be in (0, 1).")
    if minimum_separation < 0.0:
        raise ValueError("minimum_separation cannot be negative.")
    if max_iter < 1 or tolerance <= 0.0 or max_step <= 0.0:
        raise ValueError("Refinement iteration settings must be positive.")
    if damping <= 0.0 or ridge < 0.0:
        raise ValueError("damping must be positive and ridge nonnegative.")
    batch_shape = tuple(coefficients.shape[:-len(grid_shape)])
    if tuple(data.shape[:-len(grid_shape)]) != batch_shape:
        raise ValueError("data and coefficients must have the same batch shape.")
    batch = prod(batch_shape) if batch_shape else 1
    if batch > 1 and component_count is None:
        raise ValueError("Batch off-grid refinement requires a fixed component_count.")
    dims = len(grid_shape)
    # a 1-D grid runs as 1 x N sampled at 1 x M (the kernel's axis 0 then has
    # one sample and carries no position information)
    kgrid = grid_shape if dims == 2 else (1, grid_shape[0])
    ksamples = sample_shape if dims == 2 else (1, sample_shape[0])
    if not data.is_cuda:
        raise RuntimeError("off-grid refinement runs on the GPU; pass GPU tensors.")
    double = data.dtype in (torch.complex128, torch.float64)
    cplx = torch.complex128 if double else torch.complex64
    real = torch.float64 if double else torch.float32
    device = data.device
    y = data.to(dtype=cplx).reshape(batch, -1).contiguous()
    x = coefficients.to(device=device, dtype=cplx).reshape(batch, -1).contiguous()
    maximum = min(max_components, y.shape[1] - 1, x.shape[1])
    if component_count is not None:
        maximum = min(maximum, component_count)
    if maximum < 1:
        raise ValueError("At least two measurements are required.")
    if maximum > _CANDIDATES or (component_count is not None and component_count > _KMAX):
        raise NotImplementedError(
            f"off-grid refinement on MI355X takes at most {_CANDIDATES} candidate and {_KMAX} refined "
            "components per item.")
    count = component_count if component_count is not None else 0
    rows = component_count if component_count is not None else maximum
    f64 = torch.float64
    init = torch.empty((batch, rows, 2), dtype=f64, device=device)
    pos = torch.empty_like(init)
    std = torch.empty_like(init)
    amp = torch.empty((batch, rows), dtype=torch.complex128, device=device)
    norms = torch.empty((batch, 2), dtype=f64, device=device)
    counts = torch.empty((batch, 3), dtype=torch.int32, device=device)
    finfo = torch.finfo(real)
    lib = _lib.load()
    ws, ws_bytes = None, 0
    if rows > _KLDS:
        ws_bytes = int(lib.fbme_offgrid_workspace_bytes(batch))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
    with torch.cuda.device(device):
        stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
        _lib.check(lib.fbme_offgrid_refine_ws(
            stream, batch, *kgrid, *ksamples, ctypes.c_void_p(y.data_ptr()),
            ctypes.c_void_p(x.data_ptr()), 1 if double else 0, count, maximum, float(relative_threshold),
            float(minimum_separation), int(max_iter), float(tolerance), float(max_step), float(damping),
            float(ridge), float(finfo.eps), float(finfo.tiny), ctypes.c_void_p(init.data_ptr()),
            ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(amp.data_ptr()), ctypes.c_void_p(std.data_ptr()),
            ctypes.c_void_p(norms.data_ptr()), ctypes.c_void_p(counts.data_ptr()), dims,
            ctypes.c_void_p(ws.data_ptr() if ws is not None else None), ws_bytes), "fbme_offgrid_refine")
        cnt = counts.cpu()
    if bool((cnt[:, 1] < 0).any()):
        raise RuntimeError("off-grid refinement: a model order above 32 components without a workspace")
    k = int(cnt[0, 0])
    if bool((cnt[:, 0] != k).any()):
        raise RuntimeError("batch items refined to different component counts")
    out_batch = batch_shape if batch_shape else ()
    pshape = (*out_batch, k, dims)
    axes = slice(0, 2) if dims == 2 else slice(1, 2)
    return OffGridResult(
        initial_positions=init[:, :k, axes].to(real).reshape(pshape),
        positions=pos[:, :k, axes].to(real).reshape(pshape),
        amplitudes=amp[:, :k].to(cplx).reshape(*out_batch, k),
        position_std=std[:, :k, axes].to(real).reshape(pshape),
        initial_residual_norm=norms[:, 0].to(real).reshape(out_batch),
        residual_norm=norms[:, 1].to(real).reshape(out_batch),
        iterations=cnt[:, 1].to(device=device, dtype=torch.long).reshape(out_batch),
        converged=cnt[:, 2].to(device=device, dtype=torch.bool).reshape(out_batch),
    )
