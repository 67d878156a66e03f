"""Autotuner of the FFT pass shapes (plan) and the Toeplitz pass tiles (CG).

Tunables, each a pure performance choice (results are identical up to the
order of floating-point reductions inside a pass):

* plan (``FbmePlanOptions``): ``fft_kernel`` -- 0 puts every 256-point
  axis on the register 16x16 passes and the rest on the LDS Stockham passes,
  1 puts every axis on the LDS Stockham passes; ``row_lines`` -- rows per
  work item of the LDS row pass; ``col_lines`` -- columns per work item of
  the LDS column passes (the radix chain itself follows from each length).
* CG solver: ``row_tile`` -- rows per workgroup of the two E2-point row passes
  of the BTTB matvec, ``col_tile`` -- columns per workgroup of its E1-point
  column pass (``fbme_toeplitz_matvec_tiled``); keyed by the sample and
  embedding shapes (tuned on four scenes).

Every candidate is built, warmed up and timed on the device with HIP events
(plan candidates as captured hipGraph replays); the fastest wins.  Results
live in tables keyed by GPU architecture, kind and workload signature: the
table shipped beside this file (``tuning.json``, read-only) and the user's
table (``$FBME_TUNING``, default ``~/.cache/fbme_amd/tuning.json``), which
the ``tune_*`` functions update.  Lookups read the user's table first.
"""

from __future__ import annotations

import ctypes
import itertools
import json
import os
from collections import Counter
from typing import Iterable, Sequence

import torch

from . import _lib

SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
ROW_LINES = (1, 2, 4, 8, 16)
COL_LINES = (2, 4, 8, 16)
FFT_KERNELS = (0, 1)
PLAN_FIELDS = ("fft_kernel", "row_lines", "col_lines")


def user_table_path() -> str:
    return os.environ.get("FBME_TUNING") or os.path.join(os.path.expanduser("~"), ".cache", "fbme_amd",
                                                          "tuning.json")


def device_arch(device) -> str:
    props = torch.cuda.get_device_properties(torch.device(device))
    return str(getattr(props, "gcnArchName", props.name)).split(":")[0]


def plan_signature(shapes: Iterable[Sequence[int]], stable: bool) -> str:
    """'fast|256x256x160x160*512' -- profile and the multiset of scene shapes."""
    counts = Counter(tuple(int(v) for v in s) for s in shapes)
    body = ",".join(f"{'x'.join(map(str, s))}*{c}" for s, c in sorted(counts.items()))
    return ("stable|" if stable else "fast|") + body


def cg_signature(m1: int, m2: int, e1: int, e2: int) -> str:
    return f"cg|{m1}x{m2}|e{e1}x{e2}"


def _read(path: str) -> dict:
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def lookup(arch: str, signature: str) -> dict | None:
    for path in (user_table_path(), SHIPPED):
        entry = _read(path).get(arch, {}).get(signature)
        if entry is not None:
            return entry
    return None


def record(arch: str, signature: str, entry: dict, path: str | None = None) -> None:
    path = path or user_table_path()
    table = _read(path)
    table.setdefault(arch, {})[signature] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply_plan_entry(opt: _lib.FbmePlanOptions, entry: dict | None) -> _lib.FbmePlanOptions:
    if entry:
        for key in PLAN_FIELDS:
            if key in entry:
                setattr(opt, key, int(entry[key]))
    return opt


def plan_candidates(shapes: Sequence[Sequence[int]]) -> list[dict]:
    """The search space for a batch; the LDS tilings only matter for axes
    that are not on the register passes."""
    any_lds_col = any(s[0] != 256 for s in shapes)
    any_lds_row = any(s[1] != 256 for s in shapes)
    out = []
    for fk, rl, cl in itertools.product(FFT_KERNELS, ROW_LINES, COL_LINES):
        if fk == 0 and not any_lds_row and rl != ROW_LINES[0]:
            continue
        if fk == 0 and not any_lds_col and cl != COL_LINES[0]:
            continue
        out.append({"fft_kernel": fk, "row_lines": rl, "col_lines": cl})
    return out


def _synthetic_input(total: int, device) -> torch.Tensor:
    g = torch.Generator(device="cpu").manual_seed(0)
    re = torch.randn(total, generator=g)
    im = torch.randn(total, generator=g)
    return torch.complex(re, im).to(device)


def time_plan(shapes, opt, device, reps: int = 10, warmup: int = 2, y: torch.Tensor | None = None) -> float:
    """Mean milliseconds of one captured plan replay."""
    from .plan import NativePlan
    device = torch.device(device)
    with torch.cuda.device(device):
        plan = NativePlan(shapes, opt, device)
        y = _synthetic_input(plan.total_samples, device) if y is None else y
        side = torch.cuda.Stream(device=device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                plan.run(y)
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            plan.run(y)
        graph.replay()
        torch.cuda.synchronize(device)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            graph.replay()
        end.record()
        end.synchronize()
        ms = start.elapsed_time(end) / reps
        del graph, plan
    return ms


def tune_plan(shapes: Sequence[Sequence[int]], stable: bool, device, *, reps: int = 10,
              candidates: list[dict] | None = None, save: bool = True, log=None) -> dict:
    """Time every candidate on `device`; record and return the fastest
    ({fft_kernel, row_lines, col_lines, ms, default_ms, candidates})."""
    shapes = [tuple(int(v) for v in s) for s in shapes]
    base = _lib.default_options(stable)
    default = {k: int(getattr(base, k)) for k in PLAN_FIELDS}
    # the default shapes always compete, so tuning never returns a slower plan
    cands = [default] + [c for c in (candidates or plan_candidates(shapes)) if c != default]
    y = _synthetic_input(sum(s[2] * s[3] for s in shapes), device)
    timings = []
    for c in cands:
        opt = apply_plan_entry(_lib.default_options(stable), c)
        ms = time_plan(shapes, opt, device, reps=reps, y=y)
        timings.append((ms, c))
        if log:
            log(f"[tune] {c} {ms:.4f} ms")
    default_ms = timings[0][0]
    best_ms, best = min(timings, key=lambda t: t[0])
    entry = dict(best, ms=best_ms, default_ms=default_ms, candidates=len(cands))
    if save:
        record(device_arch(device), plan_signature(shapes, stable), entry)
    return entry


def tuned_plan_options(shapes, stable: bool, device, opt: _lib.FbmePlanOptions | None = None):
    """Options for a plan with the table's tuned pass shapes applied (the
    defaults when the workload was never tuned)."""
    opt = opt or _lib.default_options(stable)
    return apply_plan_entry(opt, lookup(device_arch(device), plan_signature(shapes, stable)))


def cg_tiles(device, m1, m2, e1, e2) -> tuple[int, int]:
    """(row_tile, col_tile) of the BTTB matvec; (0, 0) = the library's
    default tiles."""
    entry = lookup(device_arch(device), cg_signature(m1, m2, e1, e2))
    return (int(entry["row_tile"]), int(entry["col_tile"])) if entry else (0, 0)


def tune_cg(solver, *, reps: int = 20, save: bool = True, log=None) -> dict:
    """Time the BTTB matvec of a CG solver (torch_sbl.pcg._Solver, set up)
    over its tile space; record and return the fastest."""
    lib = _lib.load()
    t = solver.t
    m1, m2 = solver.samples
    rmax, cmax = ctypes.c_int(), ctypes.c_int()
    _lib.check(lib.fbme_toeplitz_tiles(t.e1, t.e2, ctypes.byref(rmax), ctypes.byref(cmax)), "toeplitz_tiles")
    rows = sorted({v for v in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64) if v <= rmax.value} | {rmax.value})
    cols = sorted({v for v in (2, 4, 6, 8, 12, 16) if v <= cmax.value} | {cmax.value})
    d = solver.rhs
    out = torch.empty_like(d)
    timings = []
    # (0, 0) is the library's default tiling; it always competes
    for rt, ct in [(0, 0)] + list(itertools.product(rows, cols)):
        solver.tiles = (rt, ct)
        for _ in range(2):
            solver.matvec(d, out)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            solver.matvec(d, out)
        end.record()
        end.synchronize()
        ms = start.elapsed_time(end) / reps
        timings.append((ms, rt, ct))
        if log:
            log(f"[tune] row_tile {rt} col_tile {ct} {ms:.4f} ms")
    default_ms = timings[0][0]
    best_ms, rt, ct = min(timings)
    solver.tiles = (rt, ct)
    entry = dict(row_tile=rt, col_tile=ct, ms=best_ms, default_ms=default_ms, candidates=len(timings))
    if save:
        record(device_arch(solver.dev), cg_signature(m1, m2, t.e1, t.e2), entry)
    return entry


def main(argv=None):
    """Tune the standard workloads on this GPU into a table:
    python -m fbme_amd.autotune --out fbme-sbl_amd/fbme_amd/tuning.json"""
    import argparse

    import numpy as np

    from . import scenes
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="table to write (default: the user's table)")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args(argv)
    if args.out:
        os.environ["FBME_TUNING"] = os.path.abspath(args.out)
    device = torch.device("cuda", 0)

    def log(msg):
        print(msg, flush=True)

    uneven = scenes.uneven_batch(2024, count=64)
    workloads = [("uniform 512 x 256x256/160x160", [(256, 256, 160, 160)] * 512, False),
                 ("uneven 64 scenes", [(s.grid[0], s.grid[1], s.samples[0], s.samples[1]) for s in uneven], False),
                 ("uniform 1 x 512x512/384x256 stable", [(512, 512, 384, 256)], True)]
    for name, shapes, stable in workloads:
        log(f"[tune] plan workload: {name}")
        entry = tune_plan(shapes, stable, device, reps=args.reps, log=log)
        log(f"[tune] {name}: {json.dumps(entry)}")
    from torch_sbl import SBLConfig, pcg
    for grid, samples in (((256, 256), (160, 160)), ((64, 64), (40, 40))):
        y = torch.from_numpy(np.stack([scenes.ellipse_scene(i, grid, samples, k=20).y
                                       for i in range(4)]).astype(np.complex64))
        s = pcg._Solver(y.to(device), grid, SBLConfig(mode="cg", device=str(device), hutchinson_probes=16,
                                                      auto_probe_scaling=False))
        s.init()
        s.setup_operators()
        log(f"[tune] cg matvec {grid}/{samples}")
        entry = tune_cg(s, log=log)
        log(f"[tune] cg {grid}/{samples}: {json.dumps(entry)}")


if __name__ == "__main__":
    main()
