"""HIP replacements of the reference's fused UAMP kernel wrappers.

Same function names, argument meaning, tensor contracts and error behaviour
as torch_sbl/triton_kernels.py:420-904: each wrapper validates its tensors
(GPU, exact dtype and shape, C-contiguous, same device), allocates fresh
outputs, launches on the current stream and, on failure, records the error
and re-raises.  Reductions are deterministic (fixed-order double partials)
instead of float atomics (triton_kernels.py:185-192).
"""

from __future__ import annotations

import ctypes

import torch
from torch import Tensor

from fbme_amd import _lib

_RUNTIME_ERROR: Exception | None = None

_BLOCK = 256


def triton_is_available() -> bool:
    """Whether the fused-kernel library loaded and has not failed
    (name kept from triton_kernels.py:420-427)."""
    if _RUNTIME_ERROR is not None or not torch.cuda.is_available():
        return False
    try:
        _lib.load()
    except RuntimeError:
        return False
    return True


hip_is_available = triton_is_available


def triton_error() -> Exception | None:
    """Load or most recent launch error (triton_kernels.py:430-432)."""
    if _RUNTIME_ERROR is not None:
        return _RUNTIME_ERROR
    try:
        _lib.load()
    except RuntimeError as exc:
        return exc
    return None


def reset_triton_runtime_error() -> None:
    """Clear a recorded launch failure (triton_kernels.py:440-448)."""
    global _RUNTIME_ERROR
    _RUNTIME_ERROR = None


def _record(exc: Exception) -> None:
    global _RUNTIME_ERROR
    _RUNTIME_ERROR = exc


def _validate(name, tensor, *, dtype, shape, device=None):
    if not isinstance(tensor, Tensor) or not tensor.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor.")
    if tensor.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}.")
    if tuple(tensor.shape) != tuple(shape):
        raise ValueError(
            f"{name} must have shape {tuple(shape)}, got {tuple(tensor.shape)}.")
    if not tensor.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous.")
    if device is not None and tensor.device != device:
        raise ValueError(f"{name} must be on {device}, got {tensor.device}.")


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _launch(fn, what, *args):
    try:
        _lib.check(fn(*args), what)
    except Exception as exc:
        _record(exc)
        raise


def _partials(batch, length, device, pairs=1):
    blocks = (length + _BLOCK - 1) // _BLOCK
    return torch.empty(batch * blocks * pairs, dtype=torch.float64,
                       device=device)


def uamp_message_update(y: Tensor, prediction: Tensor, old_message: Tensor,
                        data_variance: Tensor, noise_variance: Tensor, *,
                        damping: float) -> tuple[Tensor, Tensor]:
    """Fused complex UAMP message update and pseudo variance
    (triton_kernels.py:476-547)."""
    lib = _lib.load()
    if y.ndim != 2:
        raise ValueError("y must have shape [B, M].")
    b, m = y.shape
    dev = y.device
    _validate("y", y, dtype=torch.complex64, shape=(b, m))
    for name, t in (("prediction", prediction), ("old_message", old_message)):
        _validate(name, t, dtype=torch.complex64, shape=(b, m), device=dev)
    for name, t in (("data_variance", data_variance),
                    ("noise_variance", noise_variance)):
        _validate(name, t, dtype=torch.float32, shape=(b,), device=dev)
    out = torch.empty_like(old_message)
    tau = torch.empty(b, device=dev, dtype=torch.float32)
    _launch(lib.fbme_uamp_message_update, "uamp_message_update", _stream(dev),
            b, m, y.data_ptr(), prediction.data_ptr(), old_message.data_ptr(),
            data_variance.data_ptr(), noise_variance.data_ptr(),
            out.data_ptr(), tau.data_ptr(), float(damping))
    return out, tau


def uamp_denoise(prior_variance: Tensor, pseudo_variance: Tensor,
                 lifted_message: Tensor, old_mean: Tensor,
                 old_variance: Tensor, *, damping: float, tiny: float,
                 prior_floor: Tensor | None = None, hyper_shape: float = 0.0,
                 hyper_rate: float = 0.0, prior_damping: float = 0.0,
                 prior_maximum: float = 1e12):
    """Denoising, reductions and optional fused ARD update
    (triton_kernels.py:550-667)."""
    lib = _lib.load()
    if prior_variance.ndim != 2:
        raise ValueError("prior_variance must have shape [B, N].")
    b, n = prior_variance.shape
    dev = prior_variance.device
    for name, t in (("prior_variance", prior_variance),
                    ("old_variance", old_variance)):
        _validate(name, t, dtype=torch.float32, shape=(b, n), device=dev)
    _validate("pseudo_variance", pseudo_variance, dtype=torch.float32,
              shape=(b,), device=dev)
    for name, t in (("lifted_message", lifted_message), ("old_mean", old_mean)):
        _validate(name, t, dtype=torch.complex64, shape=(b, n), device=dev)
    if prior_floor is not None:
        _validate("prior_floor", prior_floor, dtype=torch.float32, shape=(b,),
                  device=dev)
    mean = torch.empty_like(old_mean)
    var = torch.empty_like(old_variance)
    out_prior = torch.empty_like(prior_variance) if prior_floor is not None else None
    dof = torch.empty(b, device=dev, dtype=torch.float32)
    vsum = torch.empty_like(dof)
    part = _partials(b, n, dev, pairs=2)
    _launch(lib.fbme_uamp_denoise, "uamp_denoise", _stream(dev), b, n,
            prior_variance.data_ptr(), pseudo_variance.data_ptr(),
            lifted_message.data_ptr(), old_mean.data_ptr(),
            old_variance.data_ptr(), mean.data_ptr(), var.data_ptr(),
            dof.data_ptr(), vsum.data_ptr(),
            prior_floor.data_ptr() if prior_floor is not None else None,
            out_prior.data_ptr() if out_prior is not None else None,
            float(damping), float(tiny), float(hyper_shape), float(hyper_rate),
            float(prior_damping), float(prior_maximum), part.data_ptr())
    return mean, var, dof, vsum, out_prior


def uamp_prior_update(mean: Tensor, posterior_variance: Tensor,
                      old_prior: Tensor, prior_floor: Tensor, *,
                      hyper_shape: float, hyper_rate: float, damping: float,
                      maximum: float) -> Tensor:
    """Independent-scene variational ARD update (triton_kernels.py:670-736)."""
    lib = _lib.load()
    if old_prior.ndim != 2:
        raise ValueError("old_prior must have shape [B, N].")
    b, n = old_prior.shape
    dev = old_prior.device
    _validate("mean", mean, dtype=torch.complex64, shape=(b, n), device=dev)
    for name, t in (("posterior_variance", posterior_variance),
                    ("old_prior", old_prior)):
        _validate(name, t, dtype=torch.float32, shape=(b, n), device=dev)
    _validate("prior_floor", prior_floor, dtype=torch.float32, shape=(b,),
              device=dev)
    out = torch.empty_like(old_prior)
    _launch(lib.fbme_uamp_prior_update, "uamp_prior_update", _stream(dev), b,
            n, mean.data_ptr(), posterior_variance.data_ptr(),
            old_prior.data_ptr(), prior_floor.data_ptr(), out.data_ptr(),
            float(hyper_shape), float(hyper_rate), float(damping),
            float(maximum))
    return out


def complex_residual_energy(data: Tensor, prediction: Tensor) -> Tensor:
    """Row reduction sum |data - prediction|^2 (triton_kernels.py:739-803).
    Always returns a tensor: the two-level reduction has no width limit."""
    lib = _lib.load()
    if data.ndim != 2:
        raise ValueError("data must have shape [B, M].")
    b, m = data.shape
    dev = data.device
    for name, t in (("data", data), ("prediction", prediction)):
        _validate(name, t, dtype=torch.complex64, shape=(b, m), device=dev)
    out = torch.empty(b, device=dev, dtype=torch.float32)
    part = _partials(b, m, dev)
    _launch(lib.fbme_complex_residual_energy, "complex_residual_energy",
            _stream(dev), b, m, data.data_ptr(), prediction.data_ptr(),
            out.data_ptr(), part.data_ptr())
    return out


def uamp_noise_update(data: Tensor, prediction: Tensor, old_noise: Tensor,
                      effective_dof: Tensor, variance_floor: Tensor, *,
                      noise_shape: float, noise_rate: float, damping: float,
                      evidence_update: bool, epsilon: float) -> Tensor:
    """Residual reduction + scalar noise update, clip and damping
    (triton_kernels.py:806-904)."""
    lib = _lib.load()
    if data.ndim != 2:
        raise ValueError("data must have shape [B, M].")
    b, m = data.shape
    dev = data.device
    for name, t in (("data", data), ("prediction", prediction)):
        _validate(name, t, dtype=torch.complex64, shape=(b, m), device=dev)
    for name, t in (("old_noise", old_noise), ("effective_dof", effective_dof),
                    ("variance_floor", variance_floor)):
        _validate(name, t, dtype=torch.float32, shape=(b,), device=dev)
    out = torch.empty_like(old_noise)
    part = _partials(b, m, dev)
    _launch(lib.fbme_uamp_noise_update, "uamp_noise_update", _stream(dev), b,
            m, data.data_ptr(), prediction.data_ptr(), old_noise.data_ptr(),
            effective_dof.data_ptr(), variance_floor.data_ptr(),
            out.data_ptr(), float(noise_shape), float(noise_rate),
            float(damping), 1 if evidence_update else 0, float(epsilon),
            part.data_ptr())
    return out
