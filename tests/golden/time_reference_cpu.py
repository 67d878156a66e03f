"""Time the reference's own CPU paths in the build container and record the
timings as a fixture (tests/golden/reference_cpu_timings.json):

    python tests/golden/time_reference_cpu.py /root/reference

  * CPU-1: torch_sbl.sbl_2d with the default SBLConfig (auto -> mode 'cg',
    max_iter 200, tol 1e-4) on the 32x32 / 24x24 scene (K = 20, 20 dB);
  * CPU-2: literature_reproduction.dense_em_sbl in complex128 on that scene
    (30 iterations);
  * CPU-3: fast_sbl under the FBME-Fast and FBME-Stable numerics on CPU
    (uamp_backend 'torch') at 256x256 / 160x160 (K = 300 ellipse, 10 dB),
    one scene per call and 8 scenes per call.

torch.set_num_threads(os.cpu_count()), one warm-up, the median of five
perf_counter runs.  The reference itself never runs on the GPU machine;
the benchmark reports these figures beside its own CPU baseline, timed on
the GPU machine's host.
"""

from __future__ import annotations

import json
import os
import platform
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "fbme-sbl_amd"))

from fbme_amd import scenes  # noqa: E402


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def median_time(fn, runs=5):
    fn()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def main(reference_root: str) -> None:
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, fast_sbl, sbl_2d
    from literature_reproduction import dense_em_sbl

    threads = os.cpu_count()
    torch.set_num_threads(threads)
    out = {"cpu_model": cpu_model(), "os_cpu_count": threads, "torch_threads": torch.get_num_threads(),
           "torch_version": torch.__version__, "method": "1 warm-up, median of 5 perf_counter runs",
           "runs": {}}
    sc = scenes.uniform_scene(1000, (32, 32), (24, 24), k=20, snr_db=20.0)
    y64 = torch.from_numpy(sc.y)
    res = {}

    def cpu1():
        res["cg"] = sbl_2d(y64, (32, 32), config=SBLConfig(device="cpu"))
    t = median_time(cpu1)
    out["runs"]["CPU-1 sbl_2d default (mode cg) 32x32/24x24"] = {
        "seconds": t, "scenes_per_s": 1.0 / t, "iterations": int(res["cg"].iterations)}

    def cpu2():
        res["dense"] = dense_em_sbl(y64.to(torch.complex128), (32, 32), iterations=30)
    t = median_time(cpu2)
    out["runs"]["CPU-2 dense_em_sbl complex128 30 iterations 32x32/24x24"] = {
        "seconds": t, "scenes_per_s": 1.0 / t}

    batch = [scenes.ellipse_scene(1000 + i, (256, 256), (160, 160), k=300, snr_db=10.0) for i in range(8)]
    y1 = torch.from_numpy(batch[0].y)
    y8 = torch.from_numpy(np.stack([s.y for s in batch]))
    cfgs = {
        "Fast": SBLConfig.paper_hybrid_uamp(
            device="cpu", uamp_backend="torch", uamp_polish_size=32, uamp_polish_iterations=2,
            uamp_polish_residual_exchange_size=0, uamp_polish_fixed_budget_diagnostics=False,
            uamp_polish_compute_truncation_diagnostic=False),
        "Stable": SBLConfig.realtime_hybrid_uamp(
            device="cpu", uamp_backend="torch", uamp_polish_size=32, uamp_polish_iterations=3,
            uamp_polish_residual_exchange_size=12, uamp_polish_fixed_budget_diagnostics=True,
            uamp_polish_compute_truncation_diagnostic=False),
    }
    for name, cfg in cfgs.items():
        t = median_time(lambda: fast_sbl(y1, (256, 256), config=cfg))
        out["runs"][f"CPU-3 fast_sbl {name} 256x256/160x160 single scene"] = {
            "seconds": t, "scenes_per_s": 1.0 / t}
        t = median_time(lambda: fast_sbl(y8, (256, 256), config=cfg))
        out["runs"][f"CPU-3 fast_sbl {name} 256x256/160x160 8 scenes per call"] = {
            "seconds": t, "scenes_per_s": 8.0 / t}
    with open(os.path.join(HERE, "reference_cpu_timings.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
