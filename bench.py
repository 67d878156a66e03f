"""Throughput of the batched FBME plan at 256x256 / 160x160 on one MI355X.

Workload (one "step"): one graph replay of the FBME-Fast plan over a batch of
B synthetic scenes (256x256 grid, 160x160 prefix, K=300 log-normal point
scatterers in a random ellipse, 10 dB SNR, polish_size 32), inputs resident
in HBM.  value = B * steps / elapsed seconds (scenes/s).

Also reported: the plain single-scene plan's scenes/s on the same GPU, the
NMSE of both on the same scenes, the roofline of the dominant kernel (the
fused row pass) timed with events on the stream it runs on, beside the
bandwidth plain streaming kernels reach on the same GPU, the float64 CPU
oracle timed on the host, a 64-scene uneven workload in one graph-captured
plan (autotuned pass shapes against the defaults), the Fast / Stable replay
latency at the shape of the reference's published RTX 4090 example (40x32
grid, 20x16 prefix) beside its published P50 / P95, and the learned warm
start and preconditioner against the plain general solver (PCG iterations,
scenes/s, NMSE).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fbme-sbl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
# sources of the row pass whose PMC traffic profiles/row_pass_traffic.json holds
ROW_PASS_SOURCES = ("fbme-sbl_amd/csrc/fbme_uamp16.hip", "fbme-sbl_amd/csrc/fbme_fft16.h",
                    "fbme-sbl_amd/csrc/fbme_device.h", "fbme-sbl_amd/csrc/fbme_types.h")


def sources_sha256(paths) -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in paths:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def archive_latency(fn, start, stop, samples=20):
    """The archive protocol (zip:examples/rtx4090_typical/PROTOCOL.md:22-33):
    `samples` event-timed calls, P50 = mean of the sorted 10th and 11th,
    P95 = the maximum (ms)."""
    import torch
    times = []
    for _ in range(samples):
        start.record()
        fn()
        stop.record()
        torch.cuda.synchronize()
        times.append(start.elapsed_time(stop))
    times.sort()
    mid = len(times) // 2
    return {"p50_ms": 0.5 * (times[mid - 1] + times[mid]), "p95_ms": times[-1], "samples": len(times)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--profile", choices=("fast", "stable"), default="fast")
    ap.add_argument("--plain-steps", type=int, default=64)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device copy (counter calibration)")
    ap.add_argument("--skip-uneven", action="store_true")
    ap.add_argument("--uneven-only", action="store_true",
                    help="time only the 64-scene uneven workload (tuned shapes) and print its JSON: the "
                         "command its kernel profile is taken with")
    ap.add_argument("--skip-pcg", action="store_true")
    ap.add_argument("--pcg-batch", type=int, default=32)
    ap.add_argument("--skip-learned", action="store_true")
    ap.add_argument("--skip-masked", action="store_true")
    ap.add_argument("--skip-joint", action="store_true")
    ap.add_argument("--skip-offgrid", action="store_true")
    ap.add_argument("--skip-dense", action="store_true")
    ap.add_argument("--learned-scenes", type=int, default=16)
    return ap.parse_args()


def pcg_kernels(device, batch=8, reps=20):
    """The PCG entry points of the general solver at 256x256/160x160 with 16
    Hutchinson probes (lanes L = 17 B): the operators are set up from the
    solver's state after one EM iteration, then each entry point is timed
    alone with events on its stream.  Algorithmic bytes per call (complex64
    = 8 B; M = M1 M2): covariance matvec on the 256 x 256 grid 24 L M +
    32 L 256 M2 + 4 B 65536 (column pass of A^H reads d, row pass with
    Gamma, column pass of A reads d again for the noise term and writes Q d;
    the [256, M2] intermediate round-trips twice; gamma once per scene), or
    through the BTTB embedding E1 x E2 16 L M + 32 L M1 E2 + 8 B E1 E2;
    low-rank preconditioner through the 256-point column transform (N1 =
    256) 24 L M + 16 L 32 M2 (v read by the projection and again by the
    correction, z written; the [32, M2] slot rows of the projection written
    and read once),
    otherwise 16 L M (factor generated from its separable phase tables);
    residual norm <r, r> 8 L M; update 48 L M; direction 24 L M; one PCG
    iteration = matvec (with <d, Q d>) + precondition (with <r, M^-1 r>) +
    update + direction + the residual norm (counted once, although only
    check iterations need it)."""
    import numpy as np
    import torch

    from fbme_amd import scenes
    from torch_sbl import SBLConfig, pcg
    grid, samples = (256, 256), (160, 160)
    sc = [scenes.ellipse_scene(1400 + i, grid, samples, k=300, snr_db=10.0) for i in range(batch)]
    y = torch.from_numpy(np.stack([s.y for s in sc]).astype(np.complex64)).to(device)
    cfg = SBLConfig(mode="cg", device=str(device), hutchinson_probes=16, auto_probe_scaling=False,
                    cg_tolerance=1e-4, max_iter=3, min_iter=1, tol=0.0)
    s = pcg._Solver(y, grid, cfg)
    s.init()
    s.posterior(1e-3)
    s.em(update=True)
    s.grid_sums(with_mean=True)
    s.setup_operators()
    s.d.copy_(s.rhs)
    s.res.copy_(s.rhs)
    lib = s.lib
    L, M = s.lanes, s.m
    m1 = samples[0]
    e1, e2 = s.t.e1, s.t.e2
    tw = s.t.tw[256] if s.grid_op else None
    np_ = s.nparts
    if s.fused:
        # the solver's own iteration on 256 x 256 grids: partial inner
        # products consumed by the update / direction passes
        mv = lambda: lib.fbme_grid_matvec_dot(  # noqa: E731
            s.stream, L, s.r, *samples, pcg._ptr(s.gamma), pcg._ptr(s.noise), pcg._ptr(s.d), pcg._ptr(s.q),
            pcg._ptr(s.gmid), pcg._ptr(tw), pcg._ptr(s.mv_part), None)
        pc = lambda: lib.fbme_lowrank_apply_fft_dot(  # noqa: E731
            s.stream, L, s.r, *samples, pcg._ptr(s.factors), pcg._ptr(s.sinv), pcg._ptr(s.bg),
            pcg._ptr(s.lr_map), pcg._ptr(s.res), pcg._ptr(s.z), pcg._ptr(s.lr_buf), pcg._ptr(s.lr_part),
            pcg._ptr(s.corr), pcg._ptr(tw), pcg._ptr(s.pc_part), None)
        up = lambda: lib.fbme_pcg_update_parts_rr(  # noqa: E731
            s.stream, L, M, pcg._ptr(s.best_x), pcg._ptr(s.z), pcg._ptr(s.d), pcg._ptr(s.q), pcg._ptr(s.rho),
            pcg._ptr(s.mv_part), np_, pcg._ptr(s.curv), pcg._ptr(s.improved), pcg._ptr(s.rr_part))
        # the solver's second half: projection, <r, z> without z, and the
        # synthesis pass writing d = z + beta d (every lane does the same
        # work whether or not its update is valid)
        pcd = lambda: lib.fbme_lowrank_apply_fft_direction(  # noqa: E731
            s.stream, L, s.r, *samples, pcg._ptr(s.factors), pcg._ptr(s.sinv), pcg._ptr(s.bg),
            pcg._ptr(s.lr_map), pcg._ptr(s.res), pcg._ptr(s.best_x), pcg._ptr(s.lr_part), pcg._ptr(s.corr),
            pcg._ptr(tw), pcg._ptr(s.rr_part), s.rr_blocks, pcg._ptr(s.rho), pcg._ptr(s.curv),
            pcg._ptr(s.nrho), pcg._ptr(s.improved))
        di = lambda: lib.fbme_pcg_direction_parts(  # noqa: E731
            s.stream, L, M, pcg._ptr(s.best_x), pcg._ptr(s.z), pcg._ptr(s.nrho), pcg._ptr(s.curv),
            pcg._ptr(s.pc_part), np_, pcg._ptr(s.lr_part), pcg._ptr(s.improved))
    # name: (call, algorithmic bytes, part of the solver iteration)
    calls = {
        "matvec": (mv if s.fused else (lambda: s.matvec(s.d, s.q, dot=s.curv)),
                   (24 * L * M + 32 * L * 256 * samples[1] + 4 * batch * 65536) if s.grid_op
                   else (16 * L * M + 32 * L * m1 * e2 + 8 * batch * e1 * e2), True),
        # the FFT form reads r twice (projection, synthesis) and writes z
        "precondition": (pc if s.fused else (lambda: s.precondition(s.res, s.z, dot=s.nrho)),
                         24 * L * M if s.lr_fft else 16 * L * M, not s.fused),
        "lane_dot": (lambda: s.lane_dot(s.res, s.res, s.rr), 8 * L * M, False),
        "update": (up if s.fused else
                   (lambda: lib.fbme_pcg_update(s.stream, L, M, pcg._ptr(s.best_x), pcg._ptr(s.z), pcg._ptr(s.d),
                                                pcg._ptr(s.q), pcg._ptr(s.rho), pcg._ptr(s.curv),
                                                pcg._ptr(s.improved))), 48 * L * M, True),
        "direction": (di if s.fused else
                      (lambda: lib.fbme_pcg_direction(s.stream, L, M, pcg._ptr(s.best_x), pcg._ptr(s.z),
                                                      pcg._ptr(s.nrho), pcg._ptr(s.nrho), pcg._ptr(s.improved),
                                                      1)), 24 * L * M, not s.fused),
    }
    if s.fused:
        # projection reads r; the synthesis reads r and d and writes d
        calls["precondition_direction"] = (pcd, 32 * L * M, True)
    s.rho.fill_(1.0)
    s.curv.fill_(2.0)
    s.nrho.fill_(1.0)
    s.improved.fill_(1)
    out = {"workload": f"{batch} scenes x 17 lanes (data + 16 probes), 256x256/160x160",
           "lanes": L, "matvec_path": "grid 256x256 (register FFT)" if s.grid_op else f"BTTB {e1}x{e2}"}
    total_ms = total_bytes = 0.0
    for name, (fn, nbytes, in_iter) in calls.items():
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            fn()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / reps
        out[name] = {"ms": ms, "bytes": nbytes, "GB/s": nbytes / (ms / 1e3) / 1e9,
                     "frac": nbytes / (ms / 1e3) / HBM_PEAK}
        # <r, r> and the separate preconditioner / direction run at check
        # iterations only; <d, Qd> and <r, z> ride in the passes
        if in_iter:
            total_ms += ms
            total_bytes += nbytes
    out["pcg_iteration"] = {"ms": total_ms, "bytes": total_bytes, "GB/s": total_bytes / (total_ms / 1e3) / 1e9,
                            "frac": total_bytes / (total_ms / 1e3) / HBM_PEAK,
                            "us_per_scene": total_ms * 1e3 / batch}
    # the solver's own iteration (first half + second half), timed as a whole
    s.active.fill_(1)
    s.rho.fill_(1.0)
    for _ in range(3):
        s._iter_first()
        s._iter_second()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        s._iter_first()
        s._iter_second()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / reps
    nb = total_bytes
    out["solver_iteration"] = {"ms": ms, "bytes": nb, "GB/s": nb / (ms / 1e3) / 1e9,
                               "frac": nb / (ms / 1e3) / HBM_PEAK, "us_per_scene": ms * 1e3 / batch,
                               "kernels": ("grid matvec + update + preconditioner/direction pass (no check)" if s.fused
                                           else "matvec + update + preconditioner + direction (no check)")}
    return out


# The reference's only published timings: one 40x32 grid / 20x16 prefix
# scene, 21 targets, 10 dB, on an RTX 4090 (archived summary.csv:8-9).
PUBLISHED_SCENE = {"grid": (40, 32), "prefix": (20, 16), "k": 21, "snr_db": 10.0,
                   "reference_rtx4090": {"fast": {"p50_ms": 0.6415, "p95_ms": 0.9513},
                                         "stable": {"p50_ms": 1.0988, "p95_ms": 1.5503}}}


def published_scene_latency(device):
    """Single-scene replay latency of the Fast and Stable plans at the shape
    of the reference's published example (synthetic scene of that shape and
    target count: the archived scene's generator and arrays are not in the
    reference), under the same protocol (resident input copied into the plan
    by replay, 20 event-timed replays, P50 / P95)."""
    import numpy as np
    import torch

    from fbme_amd import scenes
    from torch_sbl_fastest_stable import make_fastest_plan, make_stable_plan
    cfg = PUBLISHED_SCENE
    sc = scenes.uniform_scene(200, cfg["grid"], cfg["prefix"], k=cfg["k"], snr_db=cfg["snr_db"])
    data = torch.from_numpy(sc.y).to(device)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {"workload": "one scene, grid 40x32, prefix 20x16, 21 uniform targets, 10 dB (the shape of the "
                       "reference's published RTX 4090 example; synthetic data)",
           "reference_rtx4090": cfg["reference_rtx4090"]}
    for name, make in (("fast", make_fastest_plan), ("stable", make_stable_plan)):
        plan = make(data, cfg["grid"], polish_size=32, warmup=2)
        for _ in range(5):
            plan.replay(data)
        torch.cuda.synchronize()
        lat = archive_latency(lambda: plan.replay(data), start, stop)
        res = plan.replay(data, synchronize=True)
        lat["nmse_db"] = scenes.nmse_db(res.x.cpu().numpy().reshape(-1), sc.x.reshape(-1))
        lat["p50_speedup_vs_reference_rtx4090"] = cfg["reference_rtx4090"][name]["p50_ms"] / lat["p50_ms"]
        out[name] = lat
        del plan
    return out


def uneven_workload(device, reps=20, variants=("tuned", "default")):
    """64 scenes, grids from {64, 128, 192, 256}^2, per-axis sampling ratios
    in [0.3, 0.8], K in [0.5 %, 3 %] of the cells, SNR in [5, 25] dB: one
    graph-captured plan, pass shapes from the autotuner's table against the
    default shapes."""
    import numpy as np
    import torch

    from fbme_amd import _lib, scenes
    from fbme_amd.batched import UnevenPlan
    from fbme_amd.bytes_model import stagewise_scene_bytes
    batch = scenes.uneven_batch(2024, count=64)
    shapes = [(s.grid[0], s.grid[1], s.samples[0], s.samples[1]) for s in batch]
    ys = [torch.from_numpy(s.y).to(device) for s in batch]
    out = {"workload": "64 uneven scenes (grids 64^2..256^2, ratios 0.3-0.8), FBME-Fast, one plan + hipGraph"}
    for name, opt in (("tuned", None), ("default", _lib.default_options(False))):
        if name not in variants:
            continue
        plan = UnevenPlan(shapes, device=device, options=opt)
        plan.replay(ys)
        for _ in range(3):
            plan.replay()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            plan.replay()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / reps
        nm = [scenes.nmse_db(plan.scene(i)["x"].reshape(-1).cpu().numpy(), s.x.reshape(-1))
              for i, s in enumerate(batch)]
        t_us = [plan.plan.scene_info[i][2] for i in range(len(batch))]
        model = sum(stagewise_scene_bytes(a * b, c * d, t) for (a, b, c, d), t in zip(shapes, t_us))
        out[name] = {"ms_per_replay": ms, "scenes_per_s": len(batch) / (ms / 1e3),
                     "nmse_db": float(np.mean(nm)), "model_GB_per_s": model / (ms / 1e3) / 1e9,
                     "options": {"fft_kernel": plan.options.fft_kernel, "row_lines": plan.options.row_lines,
                                 "col_lines": plan.options.col_lines}}
        del plan
    return out


def masked_workload(device, batch, reps=20):
    """Observation masks (60 % of the 160 x 160 samples observed, one seeded
    mask): the batched FBME-Fast plan over the bench's scenes in one
    graph-captured replay, and the general solver's PCG iteration at 8
    scenes, each against the same work without the mask."""
    import numpy as np
    import torch

    from fbme_amd import scenes
    from fbme_amd.batched import UnevenPlan
    from torch_sbl import SBLConfig, pcg
    grid, samples = (256, 256), (160, 160)
    mask = np.random.default_rng(60).random(samples) < 0.6
    shapes = [(256, 256, 160, 160)] * len(batch)
    packed = torch.from_numpy(np.stack([np.where(mask, s.y, 0) for s in batch]).astype(np.complex64)
                              .reshape(-1)).to(device)
    out = {"workload": f"{len(batch)} scenes 256x256/160x160 K=300 10 dB, FBME-Fast, mask keeps "
                       f"{int(mask.sum())} of {mask.size} samples; PCG iteration at 8 scenes x 17 lanes"}
    for name, masks in (("masked", [mask] * len(batch)), ("unmasked", None)):
        plan = UnevenPlan(shapes, device=device, masks=masks)
        plan.replay(packed)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            plan.replay()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / reps
        nm = [scenes.nmse_db(plan.scene(i)["x"].reshape(-1).cpu().numpy(), batch[i].x.reshape(-1))
              for i in range(32)]
        out["plan_" + name] = {"ms_per_replay": ms, "scenes_per_s": len(batch) / (ms / 1e3),
                               "nmse_db_first_32": float(np.mean(nm))}
        del plan
    y = torch.from_numpy(np.stack([np.where(mask, s.y, 0) for s in batch[:8]]).astype(np.complex64)).to(device)
    cfg = SBLConfig(mode="cg", device=str(device), hutchinson_probes=16, auto_probe_scaling=False,
                    cg_tolerance=1e-4, max_iter=3, min_iter=1, tol=0.0)
    for name, mk in (("masked", torch.from_numpy(mask)), ("unmasked", None)):
        s = pcg._Solver(y, grid, cfg, None, mk)
        s.init()
        s.posterior(1e-3)
        s.em(update=True)
        s.grid_sums(with_mean=True)
        s.setup_operators()
        s.d.copy_(s.rhs)
        s.res.copy_(s.rhs)
        s.active.fill_(1)
        s.rho.fill_(1.0)
        for _ in range(3):
            s._iter_first()
            s._iter_second()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            s._iter_first()
            s._iter_second()
        stop.record()
        torch.cuda.synchronize()
        out["pcg_iteration_" + name] = {"ms": start.elapsed_time(stop) / reps}
        del s
    return out


def joint_workload(device, count=8, reps=20):
    """Joint sparsity at 256x256/160x160: `count` snapshots sharing one
    300-atom support (12 dB), 16 probes.  Shared noise makes them one PCG
    system of count + 16 lanes instead of count systems of 17 lanes; timed
    per PCG iteration and for a five-iteration solve, against the same
    snapshots solved independently."""
    import numpy as np
    import torch

    from fbme_amd import scenes
    from torch_sbl import SBLConfig, pcg
    grid, samples = (256, 256), (160, 160)
    batch = scenes.joint_scenes(1600, count, grid, samples, k=300, snr_db=12.0)
    y = torch.from_numpy(np.stack([s.y for s in batch]).astype(np.complex64)).to(device)
    kw = dict(mode="cg", device=str(device), hutchinson_probes=16, auto_probe_scaling=False,
              cg_tolerance=1e-4, max_iter=5, min_iter=1, tol=0.0)
    out = {"workload": f"{count} snapshots sharing a 300-atom support, 256x256/160x160, 12 dB, 16 probes"}
    for name, joint in (("joint_shared_noise", True), ("independent", False)):
        cfg = SBLConfig(joint_sparsity=joint, **kw)
        s = pcg._Solver(y, grid, cfg)
        s.init()
        s.posterior(1e-3)
        s.em(update=True)
        s.grid_sums(with_mean=True)
        s.setup_operators()
        s.d.copy_(s.rhs)
        s.res.copy_(s.rhs)
        s.active.fill_(1)
        s.rho.fill_(1.0)
        for _ in range(3):
            s._iter_first()
            s._iter_second()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            s._iter_first()
            s._iter_second()
        stop.record()
        torch.cuda.synchronize()
        it_ms = start.elapsed_time(stop) / reps
        lanes = s.lanes
        del s
        pcg.sbl_cg(y, grid, cfg, batch_shape=(count,))
        torch.cuda.synchronize()
        start.record()
        res = pcg.sbl_cg(y, grid, cfg, batch_shape=(count,))
        stop.record()
        torch.cuda.synchronize()
        nm = [scenes.nmse_db(res.x[i].reshape(-1).cpu().numpy(), batch[i].x.reshape(-1)) for i in range(count)]
        out[name] = {"lanes": lanes, "pcg_iteration_ms": it_ms, "solve_ms_5_em": start.elapsed_time(stop),
                     "nmse_db": float(np.mean(nm))}
    return out


def offgrid_workload(device, count=64, reps=5, cpu=True):
    """Off-grid refinement of `count` 256x256/160x160 scenes with 20 off-grid
    sources each (20 dB, fixed component count 20) in one launch, against
    the float64 oracle of the same algorithm on one scene (NumPy, host)."""
    import time

    import numpy as np
    import torch

    from fbme_amd import scenes
    from torch_sbl import refine_offgrid_fourier
    made = [scenes.offgrid_scene(2000 + i, (256, 256), (160, 160), k=20, snr_db=20.0) for i in range(count)]
    y = torch.from_numpy(np.stack([m[0].y for m in made])).to(device)
    x = torch.from_numpy(np.stack([m[0].x for m in made]).astype(np.complex64)).to(device)
    refine_offgrid_fourier(y, x, (256, 256), component_count=20)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        res = refine_offgrid_fourier(y, x, (256, 256), component_count=20)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / reps
    pos = res.positions.cpu().numpy()
    err = []
    for i, m in enumerate(made):
        d = np.abs(pos[i][:, None, :] - m[1][None, :, :])
        d = np.minimum(d, 256 - d)
        err.append(float(np.max(np.min(np.hypot(d[..., 0], d[..., 1]), axis=0))))
    out = {"workload": f"{count} scenes 256x256/160x160, 20 off-grid sources, 20 dB, component_count=20",
           "ms": ms, "scenes_per_s": count / (ms / 1e3), "max_position_error_bins": max(err),
           "mean_iterations": float(res.iterations.float().mean())}
    if cpu:
        # CPU baseline of this workload: the float64 oracle on one scene
        from oracle import offgrid_oracle as oo
        t0 = time.perf_counter()
        oo.refine_offgrid(made[0][0].y, made[0][0].x.astype(np.complex64), (256, 256), component_count=20)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / el, "unit": "scenes/s", "kind": "port",
                               "sample": "1 scene, float64 NumPy oracle of the same algorithm"}
    return out


def dense_workload(device, cpu=True):
    """mode='dense' under the dense EM configuration
    (literature_reproduction/baselines.py:254-278), 30 EM iterations:
    one 32x32/24x24 scene (M = 576), a batch of 16 such scenes, and one
    64x64/40x40 scene (M = 1600); CPU baseline: the float64 oracle on the
    32x32 scene."""
    import time

    import numpy as np
    import torch

    from fbme_amd import scenes
    from torch_sbl import SBLConfig, sbl_2d

    def cfg(iters=30):
        return SBLConfig(mode="dense", max_iter=iters, min_iter=iters, tol=0.0, hyper_shape=1e-6,
                         hyper_rate=1e-6, noise_shape=1e-6, noise_rate=1e-6, initialization="uniform",
                         damping=0.0, noise_update="evidence", active_set=False, device=str(device),
                         return_history=False, eager_change_tracking=False)

    out = {"workload": "dense EM, 30 iterations"}
    cases = (("32x32_24x24", (32, 32), (24, 24), 1), ("32x32_24x24_b16", (32, 32), (24, 24), 16),
             ("64x64_40x40", (64, 64), (40, 40), 1))
    for name, grid, samples, b in cases:
        sc = [scenes.uniform_scene(2100 + i, grid, samples, k=20, snr_db=20.0) for i in range(b)]
        y = torch.from_numpy(np.stack([s.y for s in sc]).astype(np.complex64)).to(device)
        y = y if b > 1 else y[0]
        sbl_2d(y, grid, config=cfg(2))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sbl_2d(y, grid, config=cfg())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        nm = scenes.nmse_db(res.x.reshape(b, -1)[0].cpu().numpy(), sc[0].x.reshape(-1))
        out[name] = {"s": el, "scenes_per_s": b / el, "nmse_db_first": nm}
    if cpu:
        from oracle import dense_em
        sc = scenes.uniform_scene(2100, (32, 32), (24, 24), k=20, snr_db=20.0)
        t0 = time.perf_counter()
        dense_em.dense_em_sbl(sc.y.astype(np.complex128), (32, 32), 30)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / el, "unit": "scenes/s", "kind": "port",
                               "sample": "1 scene 32x32/24x24, float64 NumPy dense EM oracle, 30 iterations"}
    return out


def learned_comparison(device, count):
    """Learned warm start and learned PCG preconditioner against the plain
    general solver (mode='cg', 16 probes, PCG tolerance 1e-4), one
    256x256/160x160 K=300 10 dB scene per solve, held-out seeds.  Kinds:
    plain (low-rank preconditioner, plain start), learned (warm start),
    learned_full (warm start + learned grid-diagonal preconditioner),
    precond (learned preconditioner, plain start)."""
    import numpy as np
    import torch

    from fbme_amd import scenes
    from torch_sbl import SBLConfig, learned, pcg
    if not os.path.exists(learned.WEIGHTS):
        return None
    net = learned.WarmStartNet.load(device)
    pnet = learned.PrecondNet.load(device) if os.path.exists(learned.PRECOND_WEIGHTS) else None
    sc = [scenes.ellipse_scene(1300 + i, (256, 256), (160, 160), k=300, snr_db=10.0) for i in range(count)]
    w = net.warm_iter
    runs = [("plain_reference_budget", "plain", 100), ("plain_target_iter", "plain", net.target_iter),
            ("plain_warm_budget", "plain", w), ("learned", "learned", w),
            (f"plain_{w - 1}", "plain", w - 1), (f"learned_{w - 1}", "learned", w - 1)]
    if pnet is not None:
        runs += [("learned_full", "learned_full", w), (f"learned_full_{w - 1}", "learned_full", w - 1),
                 ("precond_target_iter", "precond", net.target_iter), ("precond_warm_budget", "precond", w)]
    out = {"workload": f"{count} single-scene solves, 256x256/160x160, K=300, 10 dB",
           "target_iter": net.target_iter, "warm_iter": net.warm_iter}

    def solve(kind, y, cfg, batch):
        if kind == "learned":
            return learned.learned_sbl_2d(y, (256, 256), net, cfg)
        if kind == "learned_full":
            return learned.learned_sbl_2d(y, (256, 256), net, cfg, precond_net=pnet)
        if kind == "precond":
            return learned.learned_sbl_2d(y, (256, 256), None, cfg, precond_net=pnet)
        return pcg.sbl_cg(y, (256, 256), cfg, batch_shape=(batch,))

    def config(iters):
        return SBLConfig(mode="cg", device=str(device), max_iter=iters, min_iter=1, tol=0.0,
                         hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4, return_history=True)

    for name, kind, iters in runs:
        cfg = config(iters)
        solve(kind, torch.from_numpy(sc[0].y[None]).to(device), cfg, 1)
        torch.cuda.synchronize()
        secs, its, nm, fb = [], [], [], 0.0
        for s in sc:
            y = torch.from_numpy(s.y[None]).to(device)
            t0 = time.perf_counter()
            r = solve(kind, y, cfg, 1)
            torch.cuda.synchronize()
            secs.append(time.perf_counter() - t0)
            its.append(float(np.sum(r.history["cg_iterations"])))
            nm.append(scenes.nmse_db(r.x.reshape(-1).cpu().numpy(), s.x.reshape(-1)))
            fb += r.history.get("learned_fallback", [0.0])[0]
        out[name] = {"em_iterations": iters, "pcg_iterations": float(np.mean(its)),
                     "scenes_per_s": 1.0 / float(np.mean(secs)), "nmse_db": float(np.mean(nm)),
                     "median_nmse_db": float(np.median(nm)), "fallbacks": fb}
        print(f"[bench] learned comparison {name}: {out[name]}", file=sys.stderr, flush=True)
    # batched: all scenes of the sample in one solve (probes in the lane
    # dimension of every scene); a stalled learned solve reruns the batch plain
    yb = torch.from_numpy(np.stack([s.y for s in sc])).to(device)
    batched = [runs[1], runs[3]] + ([runs[6], runs[8]] if pnet is not None else []) + [runs[5]]
    for name, kind, iters in batched:
        cfg = config(iters)
        solve(kind, yb, cfg, count)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = solve(kind, yb, cfg, count)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        x = r.x.reshape(count, -1).cpu().numpy()
        nm = [scenes.nmse_db(x[i], s.x.reshape(-1)) for i, s in enumerate(sc)]
        out[name + "_batched"] = {"em_iterations": iters, "batch": count,
                                  "pcg_iterations": float(np.sum(r.history["cg_iterations"])),
                                  "scenes_per_s": count / dt, "nmse_db": float(np.mean(nm)),
                                  "fallbacks": r.history.get("learned_fallback", [0.0])[0]}
        print(f"[bench] learned comparison {name} batched: {out[name + '_batched']}", file=sys.stderr,
              flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); "
              "scenes shard across GPUs without communication, so run one "
              "process per GPU.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from fbme_amd import _lib, scenes
    from fbme_amd.bytes_model import kernel_bytes, stagewise_scene_bytes
    from torch_sbl_fastest_stable import make_fastest_plan, make_stable_plan

    device = torch.device("cuda", 0)
    if args.uneven_only:
        print(json.dumps({"uneven_workload": uneven_workload(device, args.steps, ("tuned",))}), flush=True)
        return
    grid, samples = (256, 256), (160, 160)
    n, m = grid[0] * grid[1], samples[0] * samples[1]
    b = args.batch
    make = make_fastest_plan if args.profile == "fast" else make_stable_plan
    print(f"[bench] generating {b} scenes", file=sys.stderr, flush=True)
    batch = [scenes.ellipse_scene(scenes.EVAL_SEED_BASE + i, grid, samples,
                                  k=300, snr_db=10.0) for i in range(b)]
    truth = np.stack([s.x for s in batch])
    data = torch.from_numpy(np.stack([s.y for s in batch])).to(device)

    # ---------------------------------------------------------- batched plan
    print("[bench] building batched plan", file=sys.stderr, flush=True)
    plan = make(data, grid, polish_size=32, warmup=2)
    for _ in range(args.warmup):
        plan.replay(data)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        plan.replay(data)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = start.elapsed_time(stop) / args.steps
    value = b * args.steps / (ms_per_step * args.steps / 1e3)
    x_batched = plan.result.x.detach().cpu().numpy().reshape(b, -1)
    nmse_batched = float(np.mean([scenes.nmse_db(x_batched[i], truth[i].reshape(-1))
                                  for i in range(b)]))
    batched_latency = archive_latency(lambda: plan.replay(data), start, stop)
    print(f"[bench] batched: {ms_per_step:.3f} ms/step, {value:.1f} scenes/s",
          file=sys.stderr, flush=True)

    # ------------------------------------------------------------ plain plan
    plain = None
    if not args.skip_plain:
        single = make(data[0], grid, polish_size=32, warmup=2)
        for i in range(4):
            single.replay(data[i % b])
        torch.cuda.synchronize()
        nmse_single = []
        start.record()
        for i in range(args.plain_steps):
            single.replay(data[i % b])
        stop.record()
        torch.cuda.synchronize()
        plain_ms = start.elapsed_time(stop) / args.plain_steps
        for i in range(min(b, 16)):
            r = single.replay(data[i], synchronize=True)
            nmse_single.append(scenes.nmse_db(r.x.cpu().numpy().reshape(-1),
                                              truth[i].reshape(-1)))
        nmse_batched16 = float(np.mean([scenes.nmse_db(x_batched[i], truth[i].reshape(-1))
                                        for i in range(min(b, 16))]))
        plain = dict(ms_per_scene=plain_ms, scenes_per_s=1e3 / plain_ms,
                     nmse_db=float(np.mean(nmse_single)),
                     batched_nmse_db_same_scenes=nmse_batched16,
                     latency=archive_latency(lambda: single.replay(data[0]), start, stop))
        print(f"[bench] plain single-scene: {plain_ms:.3f} ms/scene",
              file=sys.stderr, flush=True)

    # ------------------------------------------------- dominant kernel roofline
    native = plan.native_plan
    flat = plan.input_buffer.view(-1)
    kb = kernel_bytes(*grid, *samples)
    stage_ms = {}
    for name, stage in (("row", _lib.STAGE_ROW),
                        ("col_adjoint", _lib.STAGE_COL_ADJOINT),
                        ("col_forward", _lib.STAGE_COL_FORWARD),
                        ("output", _lib.STAGE_OUTPUT)):
        for _ in range(3):
            native.run_stage(stage, flat)
        torch.cuda.synchronize()
        start.record()
        for _ in range(args.kernel_reps):
            native.run_stage(stage, flat)
        stop.record()
        torch.cuda.synchronize()
        stage_ms[name] = start.elapsed_time(stop) / args.kernel_reps
    row_s = stage_ms["row"] / 1e3
    achieved = kb["row"] * b / row_s
    kname = "k_row16<2>" if native.r16_row_items > 0 else "k_row<EPT, 2>"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "row_pass_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            t = json.load(fh)
        # only a figure measured on these exact row-pass sources counts
        if (t.get("batch") == b and t.get("workload") == "256x256/160x160"
                and t.get("kernel") == kname
                and t.get("source_sha256") == sources_sha256(ROW_PASS_SOURCES)):
            traffic = t.get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                "kernel": kname + " (fused A^H row pass + denoise/ARD + A row pass)",
                "algorithmic_bytes_per_launch": kb["row"] * b,
                "avg_launch_ms": stage_ms["row"]}
    per_kernel = {k: {"ms": v, "GB/s": kb[k] * b / (v / 1e3) / 1e9}
                  for k, v in stage_ms.items()}

    # What this GPU delivers to plain streaming kernels (torch's own copy,
    # fill and sum over 1 GiB), beside the datasheet peak the fraction uses.
    def stream_gbs(fn, nbytes, reps=10):
        fn()
        torch.cuda.synchronize()
        start.record()
        for _ in range(reps):
            fn()
        stop.record()
        torch.cuda.synchronize()
        return nbytes / (start.elapsed_time(stop) / reps / 1e3) / 1e9

    import ctypes
    lib = _lib.load()
    cur = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    src = torch.empty(1 << 28, dtype=torch.float32, device=device).uniform_()
    dst = torch.empty_like(src)
    gib = src.numel() * src.element_size()
    sp, dp = ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr())
    own = {}
    for unroll in (1, 2, 4, 8):
        own[unroll] = (stream_gbs(lambda: _lib.check(lib.fbme_stream_copy(cur, sp, dp, gib, unroll), "copy"),
                                  2 * gib),
                       stream_gbs(lambda: _lib.check(lib.fbme_stream_fill(cur, dp, gib, 1.0, unroll), "fill"),
                                  gib))
    best = max(own, key=lambda k: own[k][0])
    hbm_reference = {"copy_GB_per_s": own[best][0], "write_GB_per_s": max(v[1] for v in own.values()),
                     "kernel": "k_stream_copy / k_stream_fill (libfbme_hip, one pass, 16-byte accesses; "
                               f"best of 1/2/4/8 in flight per thread: copy at {best})",
                     "copy_GB_per_s_by_unroll": {str(k): v[0] for k, v in own.items()},
                     "torch_copy_GB_per_s": stream_gbs(lambda: dst.copy_(src), 2 * gib),
                     "buffer_bytes": gib}
    roofline["frac_of_copy"] = achieved / 1e9 / hbm_reference["copy_GB_per_s"]
    del src, dst

    if args.calibrate_copy:
        # a known-size copy for the PMC calibration: 256 MiB read, 256 MiB written
        src = torch.empty(64 << 20, dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        _lib.check(lib.fbme_stream_copy(cur, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                        src.numel() * 4, 1), "calibration copy")
        torch.cuda.synchronize()

    # ----------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu_baseline and args.profile == "fast":
        from oracle import torch_cpu_plan as tcp
        threads = torch.get_num_threads()
        host = torch.from_numpy(np.stack([s.y for s in batch]))
        tcp.fast_plan(host[:8], grid)                      # warm-up
        count, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            first = count % max(b - 7, 1)
            tcp.fast_plan(host[first:first + 8], grid)
            count += 8
        el = time.perf_counter() - t0
        cpu = {"value": count / el, "unit": "scenes/s", "cores": threads, "kind": "port",
               "os_cpu_count": os.cpu_count(), "torch_threads": threads, "cpu_model": cpu_model(),
               "sample": f"{count} scenes ({el:.1f} s) of the same 256x256/160x160 workload (8 per call), FBME-Fast "
                         "restated on torch's CPU kernels (complex64 sweep, FP64 polish; "
                         "oracle/torch_cpu_plan.py), every host thread"}
        ref_path = os.path.join(ROOT, "tests", "golden", "reference_cpu_timings.json")
        if os.path.exists(ref_path):
            with open(ref_path) as fh:
                cpu["reference_in_build_container"] = json.load(fh)

    published = published_scene_latency(device)
    print(f"[bench] published-example scene: {published}", file=sys.stderr, flush=True)
    uneven = None
    if not args.skip_uneven:
        print("[bench] uneven workload", file=sys.stderr, flush=True)
        uneven = uneven_workload(device)
        print(f"[bench] uneven: {uneven}", file=sys.stderr, flush=True)
    pcg_cmp = None
    if not args.skip_pcg:
        print("[bench] pcg kernels", file=sys.stderr, flush=True)
        pcg_cmp = pcg_kernels(device, args.pcg_batch)
        print(f"[bench] pcg: {pcg_cmp}", file=sys.stderr, flush=True)
    learned_cmp = None
    if not args.skip_learned:
        learned_cmp = learned_comparison(device, args.learned_scenes)
    dense = None
    if not args.skip_dense:
        print("[bench] dense mode", file=sys.stderr, flush=True)
        dense = dense_workload(device, cpu=not args.no_cpu_baseline)
        print(f"[bench] dense: {dense}", file=sys.stderr, flush=True)
    offgrid = None
    if not args.skip_offgrid:
        print("[bench] off-grid refinement", file=sys.stderr, flush=True)
        offgrid = offgrid_workload(device, cpu=not args.no_cpu_baseline)
        print(f"[bench] offgrid: {offgrid}", file=sys.stderr, flush=True)
    joint = None
    if not args.skip_joint:
        print("[bench] joint sparsity", file=sys.stderr, flush=True)
        joint = joint_workload(device)
        print(f"[bench] joint: {joint}", file=sys.stderr, flush=True)
    masked = None
    if not args.skip_masked:
        print("[bench] masked workload", file=sys.stderr, flush=True)
        masked = masked_workload(device, batch)
        print(f"[bench] masked: {masked}", file=sys.stderr, flush=True)

    t_u = native.max_iters
    line = {
        "metric": "scenes/s at 256² at fixed NMSE, batched+learned vs plain FBME; % HBM peak",
        "value": value, "unit": "scenes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64/f32 (f64 polish)", "data": "synthetic",
        "config": {"workload": "batched FBME-%s plan (value), %d scenes/step, grid 256x256, "
                               "prefix 160x160, K=300 ellipse scatterers, 10 dB; plain single-scene "
                               "plan in plain_single_scene; learned warm start of the general solver "
                               "against the plain solver in learned_warm_start"
                               % (args.profile.capitalize(), b),
                   "batch": b, "grid": list(grid), "prefix": list(samples),
                   "polish_size": 32, "uamp_iterations": t_u},
        "roofline": roofline,
        "hbm_reference": hbm_reference,
        "cpu_baseline": cpu,
        "batched_latency": batched_latency,
        "plain_single_scene": plain,
        "published_example_scene": published,
        "speedup_vs_plain": (value / plain["scenes_per_s"]) if plain else None,
        "nmse_db_batched": nmse_batched,
        "per_kernel": per_kernel,
        "uneven_workload": uneven,
        "pcg_kernels": pcg_cmp,
        "learned_warm_start": learned_cmp,
        "masked_workload": masked,
        "joint_sparsity": joint,
        "offgrid_refinement": offgrid,
        "dense_mode": dense,
        "stagewise_model_bytes_per_scene": stagewise_scene_bytes(n, m, t_u,
                                                           args.profile == "stable"),
        "wall_s": wall,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
