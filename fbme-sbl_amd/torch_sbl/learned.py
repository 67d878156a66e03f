"""Learned warm start and learned PCG preconditioner of the general solver
(mode='cg').

Warm start: a small convolutional network maps the backprojection magnitude
|A^H y| and the solver's initial variances gamma_0 to a per-cell
log-variance correction: gamma_0 <- gamma_0 * exp(net(log|A^H y|^2,
log gamma_0)).  The corrected variances start the EM iterations.

Preconditioner: a second network of the same shape (PrecondNet) maps
|A^H y| and the current variances to the per-cell log-scale o of the
grid-diagonal background of the PCG preconditioner, P = B + F F^H with
B^-1 = A diag(w) A^H, w = exp(clamp(o)) / (N (noise + N gamma_T)) and F the
32 strongest atoms (the low-rank preconditioner of solver.py:5786-5848 with
its scalar background taken per grid cell).  P is Hermitian positive
definite for any output.  It is trained by torch_sbl.train_precond on an
unrolled-PCG loss.

The network (4 layers, 3x3, 32 channels, BF16 weights on MFMA with FP32
accumulation) runs and trains on the GPU through the fbme_cnn_* entry points
of libfbme_hip.so.  Training data are trajectories of the plain solver:
inputs from its initial state, targets from its variances after target_iter
EM iterations, so a warm-started solve skips that part of the trajectory.

The plain solver is the fallback: if a learned solve stalls (a PCG solve
exhausts cg_max_iter, the variances blow up, or the noise becomes
non-finite) the batch is solved again from the plain initialisation, and the
result says so in history['learned_fallback'].
"""

from __future__ import annotations

import ctypes
import os
from math import prod

import numpy as np
import torch

from fbme_amd import _lib

from . import pcg
from .config import SBLConfig, SBLResult

CHANNELS = (16, 32, 32, 32, 16)
CLAMP = 15.0
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warm_start.npz")
PRECOND_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "precond.npz")
# bound of the preconditioner network's per-cell log-scale correction
PRECOND_CLAMP = 4.0


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def init_params(seed: int = 0, inputs: int = 2) -> np.ndarray:
    """He-normal weights (zero on padded channels), zero biases; the output
    layer starts small so the untrained correction is close to zero."""
    rng = np.random.default_rng(seed)
    real_in = (inputs, 32, 32, 32)
    flat = []
    for l in range(4):
        ci, co = CHANNELS[l], CHANNELS[l + 1]
        w = np.zeros((9, ci, co), np.float32)
        rco = 1 if l == 3 else co
        w[:, :real_in[l], :rco] = rng.standard_normal((9, real_in[l], rco)) * np.sqrt(2.0 / (9 * real_in[l]))
        if l == 3:
            w *= 0.1
        flat += [w.reshape(-1), np.zeros(co, np.float32)]
    return np.concatenate(flat).astype(np.float32)


def unflatten(params: np.ndarray):
    """Flat parameter vector -> [(W [9, ci, co], b [co])] per layer."""
    out, o = [], 0
    for l in range(4):
        ci, co = CHANNELS[l], CHANNELS[l + 1]
        w = params[o:o + 9 * ci * co].reshape(9, ci, co)
        o += 9 * ci * co
        b = params[o:o + co]
        o += co
        out.append((w, b))
    return out


class WarmStartNet:
    """Device-resident network: fp32 parameters, bf16 forward / backward
    weights, Adam moments and shape-keyed workspaces."""

    def __init__(self, device, params: np.ndarray | None = None, seed: int = 0) -> None:
        self.lib = _lib.load()
        self.device = torch.device(device)
        n = int(self.lib.fbme_cnn_param_count())
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.size != n:
            raise ValueError(f"expected {n} parameters, got {p.size}")
        self.params = torch.from_numpy(p.copy()).to(self.device)
        nw = int(self.lib.fbme_cnn_weight_count())
        self.wf = torch.empty(nw, dtype=torch.int16, device=self.device)
        self.wb = torch.empty(nw, dtype=torch.int16, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.target_iter = 10
        self.warm_iter = 2
        self._ws = {}
        self.prepare()

    @classmethod
    def load(cls, device, path: str = WEIGHTS) -> "WarmStartNet":
        """Trained weights plus the trajectory iteration they imitate
        (target_iter) and the EM budget they were evaluated with (warm_iter)."""
        with np.load(path) as z:
            net = cls(device, z["params"])
            if "target_iter" in z:
                net.target_iter = int(z["target_iter"])
            if "warm_iter" in z:
                net.warm_iter = int(z["warm_iter"])
        return net

    def save(self, path: str, **meta) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params.cpu().numpy(), **{k: np.asarray(v) for k, v in meta.items()})

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def prepare(self) -> None:
        _lib.check(self.lib.fbme_cnn_prepare(self._stream(), _ptr(self.params), _ptr(self.wf), _ptr(self.wb)),
                   "cnn_prepare")

    def _workspace(self, b, h, w):
        key = (b, h, w)
        if key not in self._ws:
            cells = b * h * w
            parts = int(self.lib.fbme_cnn_part_floats(b, h, w))
            self._ws[key] = dict(
                act=torch.empty(3 * cells * 32, dtype=torch.int16, device=self.device),
                out=torch.empty((b, h, w), dtype=torch.float32, device=self.device),
                dz4=torch.empty(cells * 16, dtype=torch.int16, device=self.device),
                dz=torch.empty(2 * cells * 32, dtype=torch.int16, device=self.device),
                part=torch.empty(parts, dtype=torch.float32, device=self.device),
                lpart=torch.empty((cells + 255) // 256, dtype=torch.float64, device=self.device),
                loss=torch.empty(1, dtype=torch.float64, device=self.device))
        return self._ws[key]

    def features(self, bp, gamma, bpe, gsum, grid):
        b = gamma.shape[0]
        n = prod(grid)
        x = torch.empty((b, grid[0], grid[1], 16), dtype=torch.int16, device=self.device)
        _lib.check(self.lib.fbme_cnn_features(self._stream(), b, n, _ptr(bp), _ptr(gamma), _ptr(bpe), _ptr(gsum),
                                              _ptr(x)), "cnn_features")
        return x

    def forward(self, x):
        b, h, w, _ = x.shape
        ws = self._workspace(b, h, w)
        _lib.check(self.lib.fbme_cnn_forward(self._stream(), b, h, w, _ptr(x), _ptr(self.wf), _ptr(self.params),
                                             _ptr(ws["act"]), _ptr(ws["out"])), "cnn_forward")
        return ws["out"]

    def loss_and_grads(self, x, target):
        """Forward, MSE loss against target [B, H, W] and parameter gradients
        (left in self.grads); returns the loss tensor (device, FP64)."""
        b, h, w, _ = x.shape
        ws = self._workspace(b, h, w)
        out = self.forward(x)
        _lib.check(self.lib.fbme_cnn_loss(self._stream(), b * h * w, _ptr(out), _ptr(target), _ptr(ws["dz4"]),
                                          _ptr(ws["lpart"]), _ptr(ws["loss"])), "cnn_loss")
        _lib.check(self.lib.fbme_cnn_backward(self._stream(), b, h, w, _ptr(x), _ptr(ws["act"]), _ptr(self.wb),
                                              _ptr(ws["dz4"]), _ptr(ws["dz"]), _ptr(ws["part"]),
                                              _ptr(self.grads)), "cnn_backward")
        return ws["loss"]

    def train_step(self, x, target, lr: float):
        loss = self.loss_and_grads(x, target)
        self.step += 1
        _lib.check(self.lib.fbme_cnn_adam(self._stream(), self.params.numel(), _ptr(self.params), _ptr(self.grads),
                                          _ptr(self.m), _ptr(self.v), float(lr), self.step), "cnn_adam")
        self.prepare()
        return loss


class PrecondNet(WarmStartNet):
    """The same network (4 layers, 3x3, 32 channels, BF16 MFMA) emitting the
    per-cell log-scale o of the grid-diagonal preconditioner background,
    w = exp(clamp(o)) / (N (noise + N gamma_T)) (fbme_gdiag_weights).  Inputs
    (fbme_cnn_features_precond): log|A^H y|^2 and log gamma relative to their
    scene means, and log(1 + N gamma / noise) / 8.  Any output keeps the
    preconditioner Hermitian positive definite; an output of zero is the
    closed-form background."""

    def __init__(self, device, params: np.ndarray | None = None, seed: int = 0) -> None:
        super().__init__(device, init_params(seed, inputs=3) if params is None else params, seed)
        self.clamp = PRECOND_CLAMP

    @classmethod
    def load(cls, device, path: str = PRECOND_WEIGHTS) -> "PrecondNet":
        with np.load(path) as z:
            return cls(device, z["params"])

    def features(self, bp, gamma, bpe, gsum, noise, grid):
        b = gamma.shape[0]
        n = prod(grid)
        x = torch.empty((b, grid[0], grid[1], 16), dtype=torch.int16, device=self.device)
        _lib.check(self.lib.fbme_cnn_features_precond(self._stream(), b, n, _ptr(bp), _ptr(gamma), _ptr(bpe),
                                                      _ptr(gsum), _ptr(noise), _ptr(x)), "cnn_features_precond")
        return x

    def backward_from(self, x, g):
        """Parameter gradients (left in self.grads) for a loss gradient g
        [B, H, W] at the output of the last forward on x."""
        b, h, w, _ = x.shape
        ws = self._workspace(b, h, w)
        _lib.check(self.lib.fbme_cnn_output_grad(self._stream(), b * h * w, _ptr(g), _ptr(ws["dz4"])),
                   "cnn_output_grad")
        _lib.check(self.lib.fbme_cnn_backward(self._stream(), b, h, w, _ptr(x), _ptr(ws["act"]), _ptr(self.wb),
                                              _ptr(ws["dz4"]), _ptr(ws["dz"]), _ptr(ws["part"]),
                                              _ptr(self.grads)), "cnn_backward")

    def adam(self, lr: float) -> None:
        self.step += 1
        _lib.check(self.lib.fbme_cnn_adam(self._stream(), self.params.numel(), _ptr(self.params), _ptr(self.grads),
                                          _ptr(self.m), _ptr(self.v), float(lr), self.step), "cnn_adam")
        self.prepare()


def _stalled(solver, iters, cfg, sum_before) -> bool:
    sums = solver.sums.cpu().numpy()[:, 0]
    noise = solver.noise.cpu().numpy()
    blown = np.any(sums > 1e3 * np.maximum(sum_before, 1e-30))
    return bool(iters >= cfg.cg_max_iter or blown or not np.all(np.isfinite(noise)) or not np.all(np.isfinite(sums)))


def learned_sbl_2d(data, grid_shape, net: WarmStartNet | None, config: SBLConfig | None = None,
                   precond_net: PrecondNet | None = None) -> SBLResult:
    """sbl_2d(mode='cg') started from the network's variances (net) and/or
    preconditioned by the grid-diagonal background with the preconditioner
    network's correction (precond_net), with the plain solver as fallback
    (see the module docstring).  Without a config the solve runs
    net.warm_iter EM iterations (16 probes, PCG tolerance 1e-4)."""
    if net is None and precond_net is None:
        raise ValueError("learned_sbl_2d needs a warm-start or a preconditioner network")
    lead = net if net is not None else precond_net
    cfg = config or SBLConfig(mode="cg", device=str(lead.device), max_iter=lead.warm_iter, min_iter=1, tol=0.0,
                              hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4)
    y = torch.as_tensor(data)
    batch_shape = tuple(y.shape[:-2])
    b = prod(batch_shape) if batch_shape else 1
    y = y.to(device=lead.device, dtype=torch.complex64).reshape(b, *y.shape[-2:])
    grid = tuple(int(v) for v in grid_shape)
    extra = dict(precond="grid_diagonal", precond_net=precond_net) if precond_net is not None else {}

    def warm(solver):
        x = net.features(solver.bp, solver.gamma, solver.bpe, solver.sums[:, 0].contiguous(), grid)
        out = net.forward(x)
        _lib.check(net.lib.fbme_cnn_apply_warm_start(solver.stream, solver.b, solver.n, _ptr(solver.gamma),
                                                     _ptr(out), _ptr(solver.floor), CLAMP), "warm_start")
        solver.grid_sums(with_mean=False)

    try:
        res = pcg.sbl_cg(y, grid, cfg, batch_shape=batch_shape, warm_start=warm if net is not None else None,
                         stall_check=_stalled, **extra)
        res.history["learned_fallback"] = [0.0]
        return res
    except pcg.SolverStall:
        res = pcg.sbl_cg(y, grid, cfg, batch_shape=batch_shape)
        res.history["learned_fallback"] = [1.0]
        return res


def trajectory_pairs(y: torch.Tensor, grid, cfg: SBLConfig, net: WarmStartNet):
    """Run the plain solver on y [B, M1, M2]; return network inputs from its
    initial state, targets log(gamma_final / gamma_0) (clamped), the result
    and a per-scene validity mask (finite targets and noise)."""
    res = pcg.sbl_cg(y, grid, cfg, batch_shape=(y.shape[0],))
    with torch.cuda.device(y.device):
        s = pcg._Solver(y, grid, cfg)
        s.init()
        x = net.features(s.bp, s.gamma, s.bpe, s.sums[:, 0].contiguous(), grid)
        gfinal = 1.0 / res.precision.reshape(s.b, -1)
        target = torch.log(gfinal / s.gamma).clamp(-CLAMP, CLAMP).reshape(s.b, *grid).float().contiguous()
    valid = torch.isfinite(target).all(dim=(1, 2)) & torch.isfinite(res.noise_variance.reshape(-1))
    return x, torch.nan_to_num(target), res, valid
