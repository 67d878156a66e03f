"""Python handle of a native FBME plan (fbme_plan_* of libfbme_hip.so).

A :class:`NativePlan` fixes the scene shapes (uniform or uneven), owns the
device workspace and the output tensors (allocated once through PyTorch's
caching allocator), and enqueues the whole solve with one C call on the
current stream.  The reference's graph plans overwrite their outputs on every
replay (torch_sbl/cuda_graph.py:17-19); the same holds here.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from math import log
from typing import Sequence

import torch

from . import _lib

# per-scene scalar slots of the plan (FbmeScalar, csrc/fbme_types.h)
SC_NOISE = 2
SC_DOF = 4
SC_CHANGE = 8
SC_COUNT = 9


@dataclass
class PlanOutputs:
    x: torch.Tensor                    # [total_grid] complex64
    posterior_variance: torch.Tensor   # [total_grid] float32
    precision: torch.Tensor            # [total_grid] float32
    support_mask: torch.Tensor         # [total_grid] bool
    noise: torch.Tensor                # [nscenes] float32
    dof: torch.Tensor                  # [nscenes] float32
    nle: torch.Tensor                  # [nscenes] float32
    trace: torch.Tensor                # [nscenes, evaluations] float32
    exchanges: torch.Tensor            # [nscenes] int64
    diag: torch.Tensor                 # [nscenes, 4] float32
    status: torch.Tensor               # [nscenes] int32
    truncation: torch.Tensor           # [nscenes] float32 (truncation_diagnostic)


def options_from_fields(stable: bool, **fields) -> _lib.FbmePlanOptions:
    opt = _lib.default_options(stable)
    for key, value in fields.items():
        if not hasattr(opt, key):
            raise KeyError(key)
        setattr(opt, key, value)
    return opt


class NativePlan:
    """Fixed-shape native solve for a list of (N1, N2, M1, M2) scene shapes.

    ``masks`` optionally gives each scene a boolean observation mask
    [M1, M2] (None: all samples observed); the plan keeps a packed device
    copy (operators.py:57-75 restrict the operator to the observed
    samples)."""

    def __init__(self, shapes: Sequence[Sequence[int]],
                 options: _lib.FbmePlanOptions, device: torch.device,
                 masks: Sequence | None = None) -> None:
        lib = _lib.load()
        device = torch.device(device)
        if device.type != "cuda":
            raise ValueError("NativePlan needs a GPU device.")
        self.device = device
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.options = options
        flat = [v for s in self.shapes for v in s]
        arr = (ctypes.c_int * len(flat))(*flat)
        err = ctypes.create_string_buffer(512)
        packed = None
        if masks is not None and any(mk is not None for mk in masks):
            if len(masks) != len(self.shapes):
                raise ValueError("one mask (or None) per scene")
            parts, observed, group, seen = [], [], [], {}
            for i, (mk, (_, _, m1, m2)) in enumerate(zip(masks, self.shapes)):
                if mk is None:
                    mk = torch.ones((m1, m2), dtype=torch.bool)
                mk = torch.as_tensor(mk, dtype=torch.bool).cpu()
                if tuple(mk.shape) != (m1, m2):
                    raise ValueError(f"scene {i}: observation_mask must have shape {(m1, m2)}")
                count = int(mk.sum())
                if count == 0:
                    raise ValueError(f"scene {i}: observation_mask must retain at least one sample")
                parts.append(mk.reshape(-1).to(torch.uint8))
                observed.append(count)
                # scenes repeating a mask share its spectrum
                key = (self.shapes[i], mk.numpy().tobytes())
                group.append(seen.setdefault(key, i))
            if all(c == s[2] * s[3] for c, s in zip(observed, self.shapes)):
                parts = None        # every sample observed: the unmasked plan
            packed = torch.cat(parts) if parts else None
        if packed is not None:
            obs = (ctypes.c_int * len(observed))(*observed)
            grp = (ctypes.c_int * len(group))(*group)
            handle = lib.fbme_plan_create_masked(len(self.shapes), arr, obs, grp,
                                                 ctypes.byref(options), err, 512)
        else:
            handle = lib.fbme_plan_create(len(self.shapes), arr,
                                          ctypes.byref(options), err, 512)
        if not handle:
            raise ValueError(err.value.decode() or "fbme_plan_create failed")
        self._handle = handle
        self._lib = lib
        q = lambda what: int(lib.fbme_plan_query(handle, what))  # noqa: E731
        self.workspace_bytes = q(_lib.Q_WORKSPACE_BYTES)
        self.total_grid = q(_lib.Q_TOTAL_GRID)
        self.total_samples = q(_lib.Q_TOTAL_SAMPLES)
        self.max_iters = q(_lib.Q_MAX_ITERS)
        self.evaluations = q(_lib.Q_EVALUATIONS)
        self.launches = q(_lib.Q_LAUNCHES)
        self.r16_col_items = q(_lib.Q_R16_COL_ITEMS)
        self.r16_row_items = q(_lib.Q_R16_ROW_ITEMS)
        self.nscenes = len(self.shapes)
        info = (ctypes.c_longlong * 8)()
        self.scene_info = []
        for i in range(self.nscenes):
            _lib.check(lib.fbme_plan_scene_info(handle, i, info), "scene_info")
            self.scene_info.append(tuple(int(v) for v in info))
        with torch.cuda.device(device):
            self.workspace = torch.empty(self.workspace_bytes,
                                         dtype=torch.uint8, device=device)
            stream = torch.cuda.current_stream(device).cuda_stream
            _lib.check(lib.fbme_plan_bind(handle, self.workspace.data_ptr(),
                                          ctypes.c_void_p(stream)),
                       "fbme_plan_bind")
            self.mask = None
            if packed is not None:
                self.mask = packed.to(device)
                _lib.check(lib.fbme_plan_set_mask(handle, self.mask.data_ptr(),
                                                  ctypes.c_void_p(stream)),
                           "fbme_plan_set_mask")
            g, b = self.total_grid, self.nscenes
            self.out = PlanOutputs(
                x=torch.zeros(g, dtype=torch.complex64, device=device),
                posterior_variance=torch.zeros(g, dtype=torch.float32,
                                               device=device),
                precision=torch.zeros(g, dtype=torch.float32, device=device),
                support_mask=torch.zeros(g, dtype=torch.bool, device=device),
                noise=torch.zeros(b, dtype=torch.float32, device=device),
                dof=torch.zeros(b, dtype=torch.float32, device=device),
                nle=torch.zeros(b, dtype=torch.float32, device=device),
                trace=torch.zeros((b, self.evaluations), dtype=torch.float32,
                                  device=device),
                exchanges=torch.zeros(b, dtype=torch.int64, device=device),
                diag=torch.zeros((b, 4), dtype=torch.float32, device=device),
                status=torch.zeros(b, dtype=torch.int32, device=device),
                truncation=torch.zeros(b, dtype=torch.float32, device=device),
            )
        self.store_uamp = False

    # ------------------------------------------------------------------ run
    def _io(self, y: torch.Tensor) -> _lib.FbmeIO:
        o = self.out
        return _lib.FbmeIO(
            y=y.data_ptr(), x=o.x.data_ptr(),
            posterior_variance=o.posterior_variance.data_ptr(),
            precision=o.precision.data_ptr(),
            support_mask=o.support_mask.data_ptr(),
            noise=o.noise.data_ptr(), dof=o.dof.data_ptr(),
            nle=o.nle.data_ptr(), trace=o.trace.data_ptr(),
            exchanges=o.exchanges.data_ptr(), diag=o.diag.data_ptr(),
            status=o.status.data_ptr(), store_uamp=1 if self.store_uamp else 0,
            truncation=o.truncation.data_ptr())

    def run_stage(self, stage: int, y: torch.Tensor) -> None:
        """Enqueue one pass on the current plan state (kernel timing)."""
        io = self._io(y)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.fbme_plan_run_stage(self._handle, stage,
                                                 ctypes.byref(io),
                                                 ctypes.c_void_p(stream)),
                   "fbme_plan_run_stage")

    def _check_input(self, y: torch.Tensor) -> None:
        if y.device != self.device or y.dtype != torch.complex64:
            raise ValueError("y must be complex64 on the plan's device")
        if y.numel() != self.total_samples or not y.is_contiguous():
            raise ValueError(
                f"y must be contiguous with {self.total_samples} elements")

    def begin(self, y: torch.Tensor) -> None:
        """Initialisation phase of a solve driven iteration by iteration."""
        self._check_input(y)
        io = self._io(y)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.fbme_plan_begin(self._handle, ctypes.byref(io), ctypes.c_void_p(stream)),
                   "fbme_plan_begin")

    def iterate(self, y: torch.Tensor, t: int, track_change: bool) -> None:
        """UAMP iteration t (1-based); with track_change the per-scene
        relative change of the step lands in scalars()[:, SC_CHANGE]."""
        io = self._io(y)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.fbme_plan_iterate(self._handle, int(t), 1 if track_change else 0, ctypes.byref(io),
                                               ctypes.c_void_p(stream)), "fbme_plan_iterate")

    def finish(self, y: torch.Tensor) -> None:
        """Final E-step, polish (or the UAMP posterior) and outputs."""
        io = self._io(y)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.fbme_plan_finish(self._handle, ctypes.byref(io), ctypes.c_void_p(stream)),
                   "fbme_plan_finish")

    def scalars(self) -> torch.Tensor:
        """[nscenes, SC_COUNT] FP64 per-scene scalars in the plan's internal
        scene order (scenes sorted by iteration budget): power, floor,
        noise, data variance, dof, |A^H y|^2, |y|^2, sum gamma, change."""
        return self.debug_tensor(4, torch.float64, self.nscenes * SC_COUNT).view(self.nscenes, SC_COUNT)

    def run(self, y: torch.Tensor) -> PlanOutputs:
        """Enqueue one solve of packed complex64 observations on the current
        stream (no synchronisation, no allocation)."""
        self._check_input(y)
        io = self._io(y)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.fbme_plan_run(self._handle, ctypes.byref(io),
                                           ctypes.c_float(0.0),
                                           ctypes.c_void_p(stream)),
                   "fbme_plan_run")
        return self.out

    def debug_tensor(self, which: int, dtype: torch.dtype, count: int):
        """View of internal plan state (tests only): 0 gamma, 1 mean,
        2 variance, 3 backprojection, 4 scalars, 5 residual y - A mu, 6 message."""
        ptr = ctypes.c_void_p()
        _lib.check(self._lib.fbme_plan_debug_pointer(self._handle, which,
                                                      ctypes.byref(ptr)),
                   "debug_pointer")
        base = self.workspace.data_ptr()
        offset = ptr.value - base
        esize = torch.empty((), dtype=dtype).element_size()
        return self.workspace[offset: offset + count * esize].view(dtype)

    def scene_slices(self):
        """(grid offset, sample offset) of every scene in caller order."""
        return [(info[0], info[1]) for info in self.scene_info]

    def support_threshold(self, scene: int) -> float:
        n1, n2, _, _ = self.shapes[scene]
        return log(n1 * n2 / float(self.options.false_alarm))

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            try:
                self._lib.fbme_plan_destroy(handle)
            except Exception:
                pass
            self._handle = None
