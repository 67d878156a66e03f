"""Write golden input/output vectors from the reference's CPU Torch backend.

Run where the reference is available only (it never runs on the GPU machine):

    python tests/golden/make_golden.py /root/reference

What it pins (functions no closed-form check fixes end to end):
  * ``fast_sbl`` under the FBME-Fast and FBME-Stable numerics
    (SBLConfig.paper_hybrid_uamp with the profile overrides of
    torch_sbl_fastest_stable/profiles.py:63-70, 110-115), run on CPU with
    ``uamp_backend="torch"`` (the Torch UAMP branch, solver.py:2808-2850, is
    the same algebra as the Triton kernels).  Topk boundaries and the
    evidence-gate decisions depend on the whole trajectory.
  * ``sbl_2d`` with the default ``SBLConfig`` on CPU (mode "auto" -> "cg",
    solver.py:1025-1036) on the 32x32 / 24x24 workload.
  * ``dense_em_sbl`` in complex128 on the same workload.
  * ``sbl_2d`` in mode "cg" with an ``observation_mask`` (a single 32x32 /
    24x24 scene and a batch of two 64x64 / 40x40 scenes), and ``fast_sbl``
    under the Fast / Stable numerics with an ``observation_mask``.
  * ``sbl_2d`` in mode "cg" with ``joint_sparsity=True`` (shared and
    per-snapshot noise).
  * ``refine_offgrid_fourier`` on scenes with off-grid point sources, on 2-D
    and on 1-D grids, with model orders up to 48 components.
  * ``sbl_2d`` in mode "uamp" with the default ``SBLConfig`` (convergence
    checks, no polish) and with a polish (``uamp_default``).
  * options the presets leave at their defaults: a fixed noise variance,
    uniform initial variances, Hadamard probes (``options``).

Each fixture stores the scene inputs and the reference outputs as plain
arrays in a compressed ``.npz`` (no pickles).
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "fbme-sbl_amd"))

from fbme_amd import scenes  # noqa: E402


def _np(t):
    if t is None:
        return None
    return t.detach().cpu().numpy()


def _result_arrays(res, prefix=""):
    out = {}
    for name in ("x", "posterior_variance", "precision", "noise_variance",
                 "effective_dof", "negative_log_evidence", "evidence_trace",
                 "support_mask", "support_score_threshold",
                 "support_exchanges_accepted", "fixed_budget_gamma_residual",
                 "fixed_budget_noise_residual",
                 "fixed_budget_objective_decrease", "fixed_budget_monotone"):
        value = getattr(res, name, None)
        if value is not None:
            out[prefix + name] = _np(value)
    out[prefix + "iterations"] = np.array(res.iterations)
    out[prefix + "active_set_size"] = np.array(res.active_set_size)
    return out


def main(reference_root: str) -> None:
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, fast_sbl, sbl_2d
    from literature_reproduction import dense_em_sbl

    torch.set_num_threads(8)

    plan_cases = [
        ("plan_32x32_24x24", scenes.uniform_scene(1000, (32, 32), (24, 24),
                                                  k=8, snr_db=20.0)),
        ("plan_40x32_20x16", scenes.uniform_scene(1001, (40, 32), (20, 16),
                                                  k=21, snr_db=10.0)),
        ("plan_64x64_40x40", scenes.uniform_scene(1002, (64, 64), (40, 40),
                                                  k=20, snr_db=15.0)),
        ("plan_96x80_60x48", scenes.ellipse_scene(1003, (96, 80), (60, 48),
                                                  k=24, snr_db=10.0)),
    ]
    for name, scene in plan_cases:
        data = torch.from_numpy(scene.y)
        fast_cfg = SBLConfig.paper_hybrid_uamp(
            device="cpu", uamp_backend="torch",
            uamp_polish_size=32, uamp_polish_iterations=2,
            uamp_polish_residual_exchange_size=0,
            uamp_polish_fixed_budget_diagnostics=False,
            uamp_polish_compute_truncation_diagnostic=False)
        stable_cfg = SBLConfig.realtime_hybrid_uamp(
            device="cpu", uamp_backend="torch", uamp_polish_size=32,
            uamp_polish_iterations=3, uamp_polish_residual_exchange_size=12,
            uamp_polish_fixed_budget_diagnostics=True,
            uamp_polish_compute_truncation_diagnostic=False)
        arrays = dict(y=scene.y, x_true=scene.x,
                      grid=np.array(scene.grid), samples=np.array(scene.samples),
                      noise_true=np.array(scene.noise_variance))
        arrays.update(_result_arrays(fast_sbl(data, scene.grid,
                                              config=fast_cfg), "fast_"))
        arrays.update(_result_arrays(fast_sbl(data, scene.grid,
                                              config=stable_cfg), "stable_"))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name)

    scene = scenes.uniform_scene(1000, (32, 32), (24, 24), k=20, snr_db=20.0)
    data64 = torch.from_numpy(scene.y)
    data128 = data64.to(torch.complex128)
    arrays = dict(y=scene.y, x_true=scene.x, grid=np.array(scene.grid),
                  samples=np.array(scene.samples),
                  noise_true=np.array(scene.noise_variance))
    cg = sbl_2d(data64, scene.grid, config=SBLConfig(device="cpu"))
    arrays.update(_result_arrays(cg, "cg_"))
    arrays["cg_mode"] = np.array(cg.mode)
    arrays["cg_history_relative_change"] = np.array(
        cg.history.get("relative_change", []))
    arrays["cg_history_cg_iterations"] = np.array(
        cg.history.get("cg_iterations", []))
    arrays["cg_history_active_set_size"] = np.array(
        cg.history.get("active_set_size", []))
    arrays["cg_history_noise_variance"] = np.array(
        cg.history.get("noise_variance", []))
    for iters in (30, 200):
        dense = dense_em_sbl(data128, scene.grid, iterations=iters)
        arrays.update(_result_arrays(dense, f"dense{iters}_"))
    np.savez_compressed(os.path.join(HERE, "general_32x32_24x24.npz"),
                        **arrays)
    print("wrote general_32x32_24x24")


def long_cg(reference_root: str) -> None:
    """A batch of four 64x64 / 40x40 ellipse scenes (10 dB) through 100
    iterations of the CG path with 16 probes: the long-run regime where
    float32 Hutchinson estimates meet strongly pruned variances."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, sbl_2d

    torch.set_num_threads(8)
    batch = [scenes.ellipse_scene(1000 + i, (64, 64), (40, 40), k=30, snr_db=10.0)
             for i in range(4)]
    y = np.stack([s.y for s in batch]).astype(np.complex64)
    cfg = SBLConfig(device="cpu", mode="cg", hutchinson_probes=16,
                    auto_probe_scaling=False, cg_tolerance=1e-4, max_iter=100)
    res = sbl_2d(torch.from_numpy(y), (64, 64), config=cfg)
    arrays = dict(y=y, x_true=np.stack([s.x for s in batch]),
                  grid=np.array((64, 64)), samples=np.array((40, 40)))
    arrays.update(_result_arrays(res, "cg_"))
    for key in ("relative_change", "cg_iterations", "noise_variance"):
        arrays["cg_history_" + key] = np.array(res.history.get(key, []))
    np.savez_compressed(os.path.join(HERE, "general_64x64_40x40_b4.npz"), **arrays)
    print("wrote general_64x64_40x40_b4")


def observation_mask(seed: int, samples, keep: float) -> np.ndarray:
    """A seeded random observation mask over the sample block."""
    rng = np.random.default_rng(seed)
    return rng.random(samples) < keep


def masked_cg(reference_root: str) -> None:
    """``sbl_2d(mode='cg', observation_mask=...)`` on CPU: a 32x32 / 24x24
    scene with 70% of its samples observed (default configuration), and a
    batch of two 64x64 / 40x40 ellipse scenes with 60% observed (16 probes,
    tolerance 1e-4, 30 iterations)."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, sbl_2d

    torch.set_num_threads(8)
    scene = scenes.uniform_scene(1010, (32, 32), (24, 24), k=12, snr_db=20.0)
    mask = observation_mask(1010, (24, 24), 0.7)
    y = np.where(mask, scene.y, 0).astype(np.complex64)
    res = sbl_2d(torch.from_numpy(y), (32, 32), config=SBLConfig(device="cpu", mode="cg"),
                 observation_mask=torch.from_numpy(mask))
    arrays = dict(y=y, mask=mask, x_true=scene.x, grid=np.array((32, 32)),
                  samples=np.array((24, 24)))
    arrays.update(_result_arrays(res, "cg_"))
    for key in ("relative_change", "cg_iterations", "noise_variance"):
        arrays["cg_history_" + key] = np.array(res.history.get(key, []))
    np.savez_compressed(os.path.join(HERE, "masked_32x32_24x24.npz"), **arrays)
    print("wrote masked_32x32_24x24")

    batch = [scenes.ellipse_scene(1020 + i, (64, 64), (40, 40), k=30, snr_db=15.0)
             for i in range(2)]
    mask = observation_mask(1020, (40, 40), 0.6)
    y = np.stack([np.where(mask, s.y, 0) for s in batch]).astype(np.complex64)
    cfg = SBLConfig(device="cpu", mode="cg", hutchinson_probes=16,
                    auto_probe_scaling=False, cg_tolerance=1e-4, max_iter=30)
    res = sbl_2d(torch.from_numpy(y), (64, 64), config=cfg,
                 observation_mask=torch.from_numpy(mask))
    arrays = dict(y=y, mask=mask, x_true=np.stack([s.x for s in batch]),
                  grid=np.array((64, 64)), samples=np.array((40, 40)))
    arrays.update(_result_arrays(res, "cg_"))
    for key in ("relative_change", "cg_iterations", "noise_variance"):
        arrays["cg_history_" + key] = np.array(res.history.get(key, []))
    np.savez_compressed(os.path.join(HERE, "masked_64x64_40x40_b2.npz"), **arrays)
    print("wrote masked_64x64_40x40_b2")


def masked_plans(reference_root: str) -> None:
    """``fast_sbl`` under the Fast and Stable numerics with an observation
    mask: a 64x64 / 40x40 uniform scene with 70% observed and a 96x80 /
    60x48 ellipse scene with 60% observed."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, fast_sbl

    torch.set_num_threads(8)
    cases = [
        ("masked_plan_64x64_40x40", scenes.uniform_scene(1030, (64, 64), (40, 40), k=20, snr_db=15.0), 0.7),
        ("masked_plan_96x80_60x48", scenes.ellipse_scene(1031, (96, 80), (60, 48), k=24, snr_db=10.0), 0.6),
    ]
    for name, scene, keep in cases:
        mask = observation_mask(1030 + len(name), scene.samples, keep)
        y = np.where(mask, scene.y, 0).astype(np.complex64)
        fast_cfg = SBLConfig.paper_hybrid_uamp(
            device="cpu", uamp_backend="torch",
            uamp_polish_size=32, uamp_polish_iterations=2,
            uamp_polish_residual_exchange_size=0,
            uamp_polish_fixed_budget_diagnostics=False,
            uamp_polish_compute_truncation_diagnostic=False)
        stable_cfg = SBLConfig.realtime_hybrid_uamp(
            device="cpu", uamp_backend="torch", uamp_polish_size=32,
            uamp_polish_iterations=3, uamp_polish_residual_exchange_size=12,
            uamp_polish_fixed_budget_diagnostics=True,
            uamp_polish_compute_truncation_diagnostic=False)
        arrays = dict(y=y, mask=mask, x_true=scene.x, grid=np.array(scene.grid),
                      samples=np.array(scene.samples))
        m = torch.from_numpy(mask)
        arrays.update(_result_arrays(fast_sbl(torch.from_numpy(y), scene.grid, config=fast_cfg,
                                              observation_mask=m), "fast_"))
        arrays.update(_result_arrays(fast_sbl(torch.from_numpy(y), scene.grid, config=stable_cfg,
                                              observation_mask=m), "stable_"))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name)


def joint_cg(reference_root: str) -> None:
    """``sbl_2d(mode='cg', joint_sparsity=True)`` on CPU: three 32x32 / 24x24
    snapshots sharing a 12-atom support (default configuration, shared noise),
    the same with per-snapshot noise, and four 64x64 / 40x40 snapshots sharing
    30 atoms (16 probes, tolerance 1e-4, 30 iterations)."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, sbl_2d

    torch.set_num_threads(8)
    cases = [
        ("joint_32x32_24x24_b3", scenes.joint_scenes(1040, 3, (32, 32), (24, 24), k=12, snr_db=20.0),
         dict(), True),
        ("joint_32x32_24x24_b3_noise", scenes.joint_scenes(1041, 3, (32, 32), (24, 24), k=12, snr_db=15.0),
         dict(), False),
        ("joint_64x64_40x40_b4", scenes.joint_scenes(1042, 4, (64, 64), (40, 40), k=30, snr_db=10.0),
         dict(hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4, max_iter=30), True),
    ]
    for name, batch, kw, shared_noise in cases:
        y = np.stack([s.y for s in batch]).astype(np.complex64)
        cfg = SBLConfig(device="cpu", mode="cg", joint_sparsity=True, joint_noise_variance=shared_noise, **kw)
        res = sbl_2d(torch.from_numpy(y), batch[0].grid, config=cfg)
        arrays = dict(y=y, x_true=np.stack([s.x for s in batch]), grid=np.array(batch[0].grid),
                      samples=np.array(batch[0].samples), joint_noise_variance=np.array(shared_noise))
        arrays.update(_result_arrays(res, "cg_"))
        for key in ("relative_change", "cg_iterations", "noise_variance"):
            arrays["cg_history_" + key] = np.array(res.history.get(key, []))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name)


def offgrid(reference_root: str) -> None:
    """``refine_offgrid_fourier`` on CPU: 64x64 / 40x40 scenes with six
    off-grid sources (model order by BIC) in complex128 and complex64, a
    batch of two with a fixed component count, and a 128x96 / 80x60 scene
    with ten sources."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl.offgrid import refine_offgrid_fourier

    torch.set_num_threads(8)
    cases = [
        ("offgrid_64x64_40x40_c128", [1050], (64, 64), (40, 40), 6, 25.0, np.complex128, None),
        ("offgrid_64x64_40x40_c64", [1050], (64, 64), (40, 40), 6, 25.0, np.complex64, None),
        ("offgrid_64x64_40x40_b2_k5", [1051, 1052], (64, 64), (40, 40), 5, 20.0, np.complex64, 5),
        ("offgrid_128x96_80x60", [1053], (128, 96), (80, 60), 10, 15.0, np.complex64, None),
    ]
    for name, seeds, grid, samples, k, snr, dtype, count in cases:
        made = [scenes.offgrid_scene(s, grid, samples, k=k, snr_db=snr) for s in seeds]
        y = np.stack([m[0].y.astype(np.complex128) if dtype == np.complex128 else m[0].y for m in made])
        x = np.stack([m[0].x for m in made]).astype(dtype)
        res = refine_offgrid_fourier(torch.from_numpy(y.astype(dtype)), torch.from_numpy(x), grid,
                                     component_count=count)
        arrays = dict(y=y.astype(dtype), x=x, grid=np.array(grid), samples=np.array(samples),
                      true_positions=np.stack([m[1] for m in made]),
                      true_amplitudes=np.stack([m[2] for m in made]),
                      component_count=np.array(-1 if count is None else count))
        for field in ("initial_positions", "positions", "amplitudes", "position_std",
                      "initial_residual_norm", "residual_norm", "iterations", "converged"):
            arrays[field] = getattr(res, field).detach().cpu().numpy()   # leading batch dimension
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name)


OFFGRID_1D_CASES = [
    # name, seeds, N, M, sources, snr_db, dtype, component_count
    ("offgrid_1d_256_160_c64", [1054], 256, 160, 6, 25.0, np.complex64, None),
    ("offgrid_1d_512_200_c128", [1055], 512, 200, 9, 20.0, np.complex128, None),
    ("offgrid_1d_128_96_b2_k4", [1056, 1057], 128, 96, 4, 20.0, np.complex64, 4),
]


def offgrid_1d(reference_root: str) -> None:
    """``refine_offgrid_fourier`` on 1-D grids (one position coordinate per
    component: the model order's parameter count and the Fisher matrix are
    those of D = 1)."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl.offgrid import refine_offgrid_fourier

    torch.set_num_threads(8)
    for name, seeds, n, m, k, snr, dtype, count in OFFGRID_1D_CASES:
        made = [scenes.offgrid_scene_1d(s, n, m, k=k, snr_db=snr) for s in seeds]
        y = np.stack([mm[0].y for mm in made])
        x = np.stack([mm[0].x for mm in made]).astype(dtype)
        res = refine_offgrid_fourier(torch.from_numpy(y.astype(dtype)), torch.from_numpy(x), (n,),
                                     component_count=count)
        arrays = dict(y=y.astype(dtype), x=x, grid=np.array((n,)), samples=np.array((m,)),
                      true_positions=np.stack([mm[1] for mm in made]),
                      true_amplitudes=np.stack([mm[2] for mm in made]),
                      component_count=np.array(-1 if count is None else count))
        for field in ("initial_positions", "positions", "amplitudes", "position_std",
                      "initial_residual_norm", "residual_norm", "iterations", "converged"):
            arrays[field] = getattr(res, field).detach().cpu().numpy()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, "components", arrays["positions"].shape, "iterations", arrays["iterations"])


OFFGRID_WIDE_CASES = [
    # more than 32 refined components: name, seeds, grid, samples, sources, snr_db, dtype, component_count
    ("offgrid_128x128_80x80_k40", [1058], (128, 128), (80, 80), 40, 25.0, np.complex64, None),
    ("offgrid_128x128_80x80_c128_k36", [1059], (128, 128), (80, 80), 36, 25.0, np.complex128, None),
    ("offgrid_96x96_64x64_b2_c128_k48", [1060, 1061], (96, 96), (64, 64), 48, 20.0, np.complex128, 48),
    ("offgrid_1d_1024_640_c128_k40", [1062], (1024,), (640,), 40, 25.0, np.complex128, None),
]


def offgrid_wide(reference_root: str) -> None:
    """``refine_offgrid_fourier`` with model orders above 32 components
    (the default ``max_components=64``), on 2-D and 1-D grids."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl.offgrid import refine_offgrid_fourier

    torch.set_num_threads(8)
    for name, seeds, grid, samples, k, snr, dtype, count in OFFGRID_WIDE_CASES:
        if len(grid) == 2:
            made = [scenes.offgrid_scene(s, grid, samples, k=k, snr_db=snr) for s in seeds]
        else:
            made = [scenes.offgrid_scene_1d(s, grid[0], samples[0], k=k, snr_db=snr) for s in seeds]
        y = np.stack([mm[0].y for mm in made])
        x = np.stack([mm[0].x for mm in made]).astype(dtype)
        res = refine_offgrid_fourier(torch.from_numpy(y.astype(dtype)), torch.from_numpy(x), grid,
                                     component_count=count)
        arrays = dict(y=y.astype(dtype), x=x, grid=np.array(grid), samples=np.array(samples),
                      true_positions=np.stack([mm[1] for mm in made]),
                      true_amplitudes=np.stack([mm[2] for mm in made]),
                      component_count=np.array(-1 if count is None else count))
        for field in ("initial_positions", "positions", "amplitudes", "position_std",
                      "initial_residual_norm", "residual_norm", "iterations", "converged"):
            arrays[field] = getattr(res, field).detach().cpu().numpy()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, "components", arrays["positions"].shape, "iterations", arrays["iterations"],
              flush=True)


OPTION_CASES = [
    # SBLConfig options beyond the presets' defaults.  name, generator, seeds,
    # grid, samples, k, snr_db, entry ("fast" / "stable": fast_sbl under the
    # plan numerics of plan_*; "cg": sbl_2d in mode 'cg'), overrides; a
    # fixed_noise_variance of "true" is the scenes' mean true noise variance
    ("options_stable_fixed_noise_64x64_40x40", "uniform", (1100,), (64, 64), (40, 40), 20, 15.0, "stable",
     dict(fixed_noise_variance="true")),
    ("options_fast_fixed_noise_96x80_60x48", "ellipse", (1101,), (96, 80), (60, 48), 24, 10.0, "fast",
     dict(fixed_noise_variance="true")),
    ("options_fast_uniform_init_96x80_60x48", "ellipse", (1102,), (96, 80), (60, 48), 24, 10.0, "fast",
     dict(initialization="uniform")),
    ("options_cg_fixed_noise_32x32_24x24", "uniform", (1103,), (32, 32), (24, 24), 8, 20.0, "cg",
     dict(fixed_noise_variance="true", max_iter=30, return_history=True)),
    ("options_cg_joint_fixed_noise_32x32_24x24_b3", "uniform", (1104, 1105, 1106), (32, 32), (24, 24), 8, 20.0,
     "cg", dict(fixed_noise_variance="true", joint_sparsity=True, max_iter=20, return_history=True)),
    # randomised hierarchical Hadamard probes (32 of a 1024 embedding; 40 x 30
    # samples: 64 x 32 embedding, 16 probes in blocks of 8)
    ("options_cg_hadamard_32x32_24x24", "uniform", (1107,), (32, 32), (24, 24), 8, 20.0, "cg",
     dict(probe_type="hadamard", max_iter=30, return_history=True)),
    ("options_cg_hadamard_48x40_40x30", "uniform", (1108,), (48, 40), (40, 30), 10, 15.0, "cg",
     dict(probe_type="hadamard", hutchinson_probes=16, auto_probe_scaling=False, max_iter=25,
          return_history=True)),
    # Q d through full-grid FFTs instead of the BTTB embedding (the same operator)
    ("options_cg_fourier_32x32_24x24", "uniform", (1109,), (32, 32), (24, 24), 8, 20.0, "cg",
     dict(cg_matvec="fourier", max_iter=30, return_history=True)),
    # low-rank preconditioners of rank 8 and 0 (the scalar background)
    ("options_cg_rank8_48x40_40x30", "uniform", (1110,), (48, 40), (40, 30), 10, 15.0, "cg",
     dict(cg_preconditioner_rank=8, max_iter=25, return_history=True)),
    ("options_cg_rank0_32x32_24x24", "uniform", (1111,), (32, 32), (24, 24), 8, 20.0, "cg",
     dict(cg_preconditioner_rank=0, max_iter=20, return_history=True)),
    # the randomised Nystrom preconditioner (rank 32 and 12; shared joint noise)
    ("options_cg_nystrom_32x32_24x24", "uniform", (1112,), (32, 32), (24, 24), 8, 20.0, "cg",
     dict(cg_preconditioner="nystrom", max_iter=30, return_history=True)),
    ("options_cg_nystrom_rank12_64x64_40x40", "uniform", (1117,), (64, 64), (40, 40), 20, 15.0, "cg",
     dict(cg_preconditioner="nystrom", cg_preconditioner_rank=12, max_iter=20, return_history=True)),
    ("options_cg_joint_nystrom_32x32_24x24_b3", "uniform", (1114, 1115, 1116), (32, 32), (24, 24), 8, 20.0,
     "cg", dict(cg_preconditioner="nystrom", joint_sparsity=True, max_iter=15, return_history=True)),
    # adaptive probe counts: 16, then 24 (iterations 6-15), then all 48 / 32
    ("options_cg_adaptive_probes_64x64_40x40", "uniform", (1118,), (64, 64), (40, 40), 20, 15.0, "cg",
     dict(adaptive_probes=True, hutchinson_probes=48, auto_probe_scaling=False, max_iter=25,
          return_history=True)),
    ("options_cg_joint_adaptive_probes_32x32_24x24_b2", "uniform", (1119, 1120), (32, 32), (24, 24), 8, 20.0,
     "cg", dict(adaptive_probes=True, joint_sparsity=True, max_iter=20, return_history=True)),
    # pattern-coupled prior with a fixed coupling, zero and circular boundaries
    ("options_cg_spatial_zero_32x32_24x24", "ellipse", (1121,), (32, 32), (24, 24), 30, 15.0, "cg",
     dict(spatial_coupling=0.3, max_iter=25, return_history=True)),
    ("options_cg_spatial_circular_48x40_40x30", "ellipse", (1122,), (48, 40), (40, 30), 40, 15.0, "cg",
     dict(spatial_coupling=0.2, spatial_boundary="circular", max_iter=20, return_history=True)),
]


def options(reference_root: str) -> None:
    """Reference runs with SBLConfig options the presets leave at their
    defaults: a fixed noise variance (plan numerics and mode 'cg', also with
    joint sparsity), uniform initial variances and Hadamard probes."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, fast_sbl, sbl_2d

    torch.set_num_threads(8)
    plan_over = dict(
        fast=dict(uamp_polish_size=32, uamp_polish_iterations=2, uamp_polish_residual_exchange_size=0,
                  uamp_polish_fixed_budget_diagnostics=False, uamp_polish_compute_truncation_diagnostic=False),
        stable=dict(uamp_polish_size=32, uamp_polish_iterations=3, uamp_polish_residual_exchange_size=12,
                    uamp_polish_fixed_budget_diagnostics=True, uamp_polish_compute_truncation_diagnostic=False))
    for name, gen, seeds, grid, samples, k, snr, entry, over in OPTION_CASES:
        fn = dict(uniform=scenes.uniform_scene, ellipse=scenes.ellipse_scene)[gen]
        batch = [fn(seed, grid, samples, k=k, snr_db=snr) for seed in seeds]
        y = np.stack([sc.y for sc in batch]).astype(np.complex64)
        over = dict(over)
        if over.get("fixed_noise_variance") == "true":
            over["fixed_noise_variance"] = float(np.float32(np.mean([sc.noise_variance for sc in batch])))
        data = torch.from_numpy(y if len(seeds) > 1 else y[0])
        if entry == "cg":
            res = sbl_2d(data, grid, config=SBLConfig(mode="cg", device="cpu", **over))
        else:
            cfg = SBLConfig.paper_hybrid_uamp(device="cpu", uamp_backend="torch", **plan_over[entry], **over)
            res = fast_sbl(data, grid, config=cfg)
        arrays = dict(y=y, x_true=np.stack([sc.x for sc in batch]), grid=np.array(grid), samples=np.array(samples),
                      entry=np.array(entry))
        arrays.update(_result_arrays(res))
        for key in ("relative_change", "cg_iterations", "noise_variance", "hutchinson_probes"):
            if res.history.get(key):
                arrays["history_" + key] = np.array(res.history[key], np.float64)
        for key, value in over.items():
            arrays["override_" + key] = np.array(value)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, "iterations", res.iterations, "noise", _np(res.noise_variance).reshape(-1)[:3],
              "active", res.active_set_size, flush=True)


HEADLINE_CASES = [
    # name, generator, seed, grid, samples, k, snr_db, profiles
    ("headline_256x256_160x160_s1000", "ellipse", 1000, (256, 256), (160, 160), 300, 10.0, ("fast", "stable")),
    ("headline_256x256_160x160_s1001", "ellipse", 1001, (256, 256), (160, 160), 300, 10.0, ("fast", "stable")),
    ("headline_256x256_160x160_s1002", "ellipse", 1002, (256, 256), (160, 160), 300, 10.0, ("fast", "stable")),
    ("headline_512x512_384x256_0db", "clustered", 1000, (512, 512), (384, 256), 1200, 0.0, ("stable",)),
    ("headline_512x512_384x256_10db", "clustered", 1000, (512, 512), (384, 256), 1200, 10.0, ("stable",)),
    ("headline_512x512_384x256_20db", "clustered", 1000, (512, 512), (384, 256), 1200, 20.0, ("stable",)),
]


def headline_scene(generator, seed, grid, samples, k, snr_db):
    """The scene of one headline fixture (regenerated by the tests from its
    seed; the fixture pins the bytes of ``y`` by their SHA-256)."""
    fn = scenes.ellipse_scene if generator == "ellipse" else scenes.clustered_scene
    return fn(seed, grid, samples, k=k, snr_db=snr_db)


def y_digest(y: np.ndarray) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(y, dtype=np.complex64).tobytes()).hexdigest()


def headline(reference_root: str) -> None:
    """``fast_sbl`` under the FBME-Fast / FBME-Stable numerics at the
    headline sizes, where K >> polish_size: 256x256 / 160x160 with a
    300-scatterer ellipse at 10 dB (three seeds, both profiles), and
    512x512 / 384x256 with 1200 clustered scatterers at 0 / 10 / 20 dB
    (Stable).  The inputs are not stored: the fixture keeps their SHA-256
    and the generator arguments.  ``x`` is exactly zero outside the polished
    atoms (prune_tail), so it is stored sparsely; the posterior variance is
    stored on the polished atoms plus its grid sum."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, fast_sbl

    torch.set_num_threads(8)
    fast_cfg = SBLConfig.paper_hybrid_uamp(
        device="cpu", uamp_backend="torch",
        uamp_polish_size=32, uamp_polish_iterations=2,
        uamp_polish_residual_exchange_size=0,
        uamp_polish_fixed_budget_diagnostics=False,
        uamp_polish_compute_truncation_diagnostic=False)
    stable_cfg = SBLConfig.realtime_hybrid_uamp(
        device="cpu", uamp_backend="torch", uamp_polish_size=32,
        uamp_polish_iterations=3, uamp_polish_residual_exchange_size=12,
        uamp_polish_fixed_budget_diagnostics=True,
        uamp_polish_compute_truncation_diagnostic=False)
    for name, gen, seed, grid, samples, k, snr, profiles in HEADLINE_CASES:
        scene = headline_scene(gen, seed, grid, samples, k, snr)
        arrays = dict(generator=np.array(gen), seed=np.array(seed), grid=np.array(grid),
                      samples=np.array(samples), k=np.array(k), snr_db=np.array(snr),
                      y_sha256=np.array(y_digest(scene.y)))
        data = torch.from_numpy(scene.y)
        for prof in profiles:
            res = fast_sbl(data, grid, config=fast_cfg if prof == "fast" else stable_cfg)
            x = _np(res.x).reshape(-1)
            nz = np.flatnonzero(x)
            var = _np(res.posterior_variance).reshape(-1)
            p = prof + "_"
            arrays[p + "x_index"] = nz.astype(np.int64)
            arrays[p + "x_value"] = x[nz]
            arrays[p + "variance_on_support"] = var[nz]
            arrays[p + "variance_sum"] = np.array(float(np.sum(var, dtype=np.float64)))
            arrays[p + "support_mask_index"] = np.flatnonzero(_np(res.support_mask).reshape(-1)).astype(np.int64)
            for field in ("noise_variance", "effective_dof", "negative_log_evidence",
                          "evidence_trace", "support_exchanges_accepted",
                          "fixed_budget_gamma_residual", "fixed_budget_noise_residual",
                          "fixed_budget_monotone"):
                value = getattr(res, field, None)
                if value is not None:
                    arrays[p + field] = _np(value)
            arrays[p + "iterations"] = np.array(res.iterations)
            arrays[p + "active_set_size"] = np.array(res.active_set_size)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, flush=True)


GENERIC_CASES = [
    # name, generator, seed, grid, samples, k, snr_db, preset, overrides
    ("generic_hybrid_64x64_40x40", "uniform", 1060, (64, 64), (40, 40), 20, 15.0, "fast_hybrid_uamp", {}),
    ("generic_hybrid_96x80_60x48", "ellipse", 1061, (96, 80), (60, 48), 24, 10.0, "fast_hybrid_uamp",
     dict(uamp_polish_max_truncation_delta=50.0)),
    ("generic_hybrid_256x256_160x160", "ellipse", 1000, (256, 256), (160, 160), 300, 10.0, "fast_hybrid_uamp",
     dict(uamp_polish_max_truncation_delta=1e9)),
    ("generic_uamp_64x64_40x40", "uniform", 1060, (64, 64), (40, 40), 20, 15.0, "fast_uamp", {}),
    ("generic_uamp_96x80_60x48", "ellipse", 1061, (96, 80), (60, 48), 24, 10.0, "fast_uamp", {}),
    ("generic_uamp_256x256_160x160", "ellipse", 1000, (256, 256), (160, 160), 300, 10.0, "fast_uamp", {}),
]


def generic(reference_root: str) -> None:
    """``fast_sbl`` under the public presets the generic graph plan takes:
    ``SBLConfig.fast_hybrid_uamp`` (the default of make_cuda_graph_sbl_plan,
    with the truncation diagnostic on) and ``SBLConfig.fast_uamp`` (one
    reduced evaluation with diffuse tail loading: no pruning,
    re-initialisation or evidence gate).  The inputs are regenerated from the
    stored generator arguments (SHA-256 of y kept); x is exactly zero off the
    polished atoms, so it is stored sparsely, and the posterior variance /
    precision are stored on the polished atoms plus their grid sums."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, fast_sbl

    torch.set_num_threads(8)
    for name, gen, seed, grid, samples, k, snr, preset, over in GENERIC_CASES:
        fn = dict(uniform=scenes.uniform_scene, ellipse=scenes.ellipse_scene)[gen]
        scene = fn(seed, grid, samples, k=k, snr_db=snr)
        cfg = getattr(SBLConfig, preset)(device="cpu", uamp_backend="torch", **over)
        res = fast_sbl(torch.from_numpy(scene.y), grid, config=cfg)
        x = _np(res.x).reshape(-1)
        nz = np.flatnonzero(x)
        var = _np(res.posterior_variance).reshape(-1)
        prec = _np(res.precision).reshape(-1)
        arrays = dict(generator=np.array(gen), seed=np.array(seed), grid=np.array(grid),
                      samples=np.array(samples), k=np.array(k), snr_db=np.array(snr),
                      preset=np.array(preset), y_sha256=np.array(y_digest(scene.y)),
                      x_index=nz.astype(np.int64), x_value=x[nz], variance_on_support=var[nz],
                      variance_sum=np.array(float(np.sum(var, dtype=np.float64))),
                      precision_on_support=prec[nz],
                      precision_sum=np.array(float(np.sum(prec, dtype=np.float64))),
                      iterations=np.array(res.iterations), active_set_size=np.array(res.active_set_size))
        for key, value in over.items():
            arrays["override_" + key] = np.array(value)
        for field in ("noise_variance", "effective_dof", "negative_log_evidence", "evidence_trace",
                      "truncation_delta_bound", "hard_pruning_accepted", "truncation_certified"):
            value = getattr(res, field, None)
            if value is not None:
                arrays[field] = _np(value)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, {f: arrays.get(f) for f in ("truncation_delta_bound", "hard_pruning_accepted")},
              "support", nz.size, flush=True)


ACTIVE_CASES = [
    # name, seed, grid, samples, k, snr_db, SBLConfig overrides: the reference
    # switches to 58, 161 and 37 atoms; then max_size 16 and min_size 48
    ("active_64x64_40x40_k60", 1080, (64, 64), (40, 40), 60, 20.0, dict(max_iter=60)),
    ("active_48x48_32x32_k50", 1082, (48, 48), (32, 32), 50, 20.0, dict(max_iter=60)),
    ("active_64x64_40x40_k45", 1083, (64, 64), (40, 40), 45, 30.0, dict(max_iter=60)),
    ("active_64x64_40x40_k60_max16", 1080, (64, 64), (40, 40), 60, 20.0,
     dict(max_iter=40, active_set_max_size=16, active_set_min_size=8)),
    ("active_64x64_40x40_k45_min48", 1083, (64, 64), (40, 40), 45, 30.0,
     dict(max_iter=40, active_set_min_size=48)),
]


def active_set(reference_root: str) -> None:
    """``sbl_2d`` in mode 'cg' on single scenes where the reference's
    active-set switch (solver.py:3184-3216, 1316-1341) takes more than 32
    atoms, and with non-default active_set_max_size / active_set_min_size."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, sbl_2d

    torch.set_num_threads(8)
    for name, seed, grid, samples, k, snr, over in ACTIVE_CASES:
        scene = scenes.uniform_scene(seed, grid, samples, k=k, snr_db=snr)
        res = sbl_2d(torch.from_numpy(scene.y), grid, config=SBLConfig(device="cpu", mode="cg", **over))
        arrays = dict(y=scene.y, x_true=scene.x, grid=np.array(grid), samples=np.array(samples))
        for key, value in over.items():
            arrays["override_" + key] = np.array(value)
        arrays.update(_result_arrays(res, "cg_"))
        for key in ("relative_change", "cg_iterations", "noise_variance", "active_set_size"):
            arrays["cg_history_" + key] = np.array(res.history.get(key, []))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, "active set", res.active_set_size, "iterations", res.iterations, flush=True)


PRECOND_CASES = [
    # name, seed, grid, samples, k, snr_db, mask keep fraction, SBLConfig overrides
    ("precond_circulant_32x32_24x24", 1000, (32, 32), (24, 24), 20, 20.0, None,
     dict(cg_preconditioner="circulant", max_iter=40)),
    # the reference's circulant PCG reaches its 150-iteration cap after a
    # few EM iterations on 64 x 64 grids: five iterations, 0 dB
    ("precond_circulant_64x64_40x40", 1100, (64, 64), (40, 40), 30, 0.0, None,
     dict(cg_preconditioner="circulant", hutchinson_probes=16, auto_probe_scaling=False, cg_tolerance=1e-4,
          max_iter=5)),
    ("precond_circulant_masked_64x64_40x40", 1101, (64, 64), (40, 40), 30, 0.0, 0.7,
     dict(cg_preconditioner="circulant", hutchinson_probes=16, auto_probe_scaling=False, max_iter=5)),
    ("precond_none_32x32_24x24", 1000, (32, 32), (24, 24), 20, 20.0, None,
     dict(cg_preconditioner="none", max_iter=40)),
]


def dense_joint(reference_root: str) -> None:
    """``sbl_2d(mode='dense', joint_sparsity=True)``: three 32x32 / 24x24
    snapshots sharing a 12-atom support, shared and per-snapshot noise,
    complex64 (30 iterations)."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, sbl_2d

    torch.set_num_threads(8)
    for name, seed, shared in (("dense_joint_32x32_24x24_b3", 1040, True),
                               ("dense_joint_32x32_24x24_b3_noise", 1041, False)):
        batch = scenes.joint_scenes(seed, 3, (32, 32), (24, 24), k=12, snr_db=20.0)
        y = np.stack([s.y for s in batch]).astype(np.complex64)
        cfg = SBLConfig(device="cpu", mode="dense", joint_sparsity=True, joint_noise_variance=shared, max_iter=30)
        res = sbl_2d(torch.from_numpy(y), (32, 32), config=cfg)
        arrays = dict(y=y, x_true=np.stack([s.x for s in batch]), grid=np.array((32, 32)),
                      samples=np.array((24, 24)), joint_noise_variance=np.array(shared))
        arrays.update(_result_arrays(res, "dense_"))
        for key in ("relative_change", "noise_variance"):
            arrays["dense_history_" + key] = np.array(res.history.get(key, []))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, "iterations", res.iterations, flush=True)


def preconditioners(reference_root: str) -> None:
    """``sbl_2d`` in mode 'cg' with cg_preconditioner='circulant' (the Strang
    BCCB preconditioner, operators.py:373-416) and 'none'."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, sbl_2d

    torch.set_num_threads(8)
    for name, seed, grid, samples, k, snr, keep, over in PRECOND_CASES:
        scene = scenes.uniform_scene(seed, grid, samples, k=k, snr_db=snr)
        mask = observation_mask(seed, samples, keep) if keep else None
        y = scene.y if mask is None else np.where(mask, scene.y, 0).astype(np.complex64)
        res = sbl_2d(torch.from_numpy(y), grid, config=SBLConfig(device="cpu", mode="cg", **over),
                     observation_mask=None if mask is None else torch.from_numpy(mask))
        arrays = dict(y=y, x_true=scene.x, grid=np.array(grid), samples=np.array(samples))
        if mask is not None:
            arrays["mask"] = mask
        for key, value in over.items():
            arrays["override_" + key] = np.array(value)
        arrays.update(_result_arrays(res, "cg_"))
        for key in ("relative_change", "cg_iterations", "noise_variance", "active_set_size"):
            arrays["cg_history_" + key] = np.array(res.history.get(key, []))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, "iterations", res.iterations, "pcg", res.history["cg_iterations"][:6], flush=True)


UAMP_CASES = [
    # name, generator, seeds (one scene per seed), grid, samples, k, snr_db,
    # SBLConfig overrides of mode='uamp' (tol 1e-4, no polish by default)
    ("uamp_64x64_40x40", "uniform", (1090,), (64, 64), (40, 40), 15, 20.0, {}),
    ("uamp_128x96_48x40_b2", "ellipse", (1091, 1092), (128, 96), (48, 40), 30, 15.0, {}),
    # converges before its budget of 42 iterations (stops at the check of iteration 25)
    ("uamp_64x48_60x44_tol1e-3", "uniform", (1099,), (64, 48), (60, 44), 5, 30.0,
     dict(max_iter=60, uamp_iteration_offset=40.0, tol=1e-3)),
    ("uamp_64x64_40x40_polish", "uniform", (1094,), (64, 64), (40, 40), 12, 20.0, dict(uamp_polish_size=16)),
    # uniform initial variances; a fixed budget T_u = max_iter
    ("uamp_64x64_40x40_uniform", "uniform", (1095,), (64, 64), (40, 40), 15, 20.0, dict(initialization="uniform")),
    ("uamp_96x64_48x40_fixed_budget", "ellipse", (1096,), (96, 64), (48, 40), 20, 15.0,
     dict(uamp_auto_iterations=False, max_iter=24)),
]


def uamp_default(reference_root: str) -> None:
    """``sbl_2d`` in mode 'uamp' with the SBLConfig defaults (adaptive damping,
    T_u = ceil(10 + 1.1 N / M), convergence checks at tol 1e-4 from iteration
    10 every 2 iterations with patience 2, no polish), on CPU with the Torch
    UAMP branch, with the per-iteration history; one case converges before
    its iteration budget and one adds a 16-atom polish."""
    sys.path.insert(0, reference_root)
    import torch
    from torch_sbl import SBLConfig, sbl_2d

    torch.set_num_threads(8)
    for name, gen, seeds, grid, samples, k, snr, over in UAMP_CASES:
        fn = dict(uniform=scenes.uniform_scene, ellipse=scenes.ellipse_scene)[gen]
        batch = [fn(seed, grid, samples, k=k, snr_db=snr) for seed in seeds]
        y = np.stack([sc.y for sc in batch]).astype(np.complex64)
        cfg = SBLConfig(mode="uamp", device="cpu", uamp_backend="torch", return_history=True, **over)
        res = sbl_2d(torch.from_numpy(y), grid, config=cfg)
        arrays = dict(y=y, truth=np.stack([sc.x for sc in batch]), grid=np.array(grid),
                      samples=np.array(samples), converged=np.array(bool(res.converged)),
                      relative_change=np.array(float(res.relative_change)),
                      history_relative_change=np.array(res.history["relative_change"], np.float64),
                      history_noise_variance=np.array(res.history["noise_variance"], np.float64),
                      history_effective_dof=np.array(res.history["effective_dof"], np.float64))
        arrays.update(_result_arrays(res))
        for key, value in over.items():
            arrays["override_" + key] = np.array(value)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name, "iterations", res.iterations, "converged", res.converged,
              "change", res.relative_change, flush=True)


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    which = sys.argv[2] if len(sys.argv) > 2 else "all"
    if which == "uamp":
        uamp_default(root)
    elif which == "offgrid_1d":
        offgrid_1d(root)
    elif which == "offgrid_wide":
        offgrid_wide(root)
    elif which == "options":
        options(root)
    elif which == "dense_joint":
        dense_joint(root)
    elif which == "precond":
        preconditioners(root)
    elif which == "active":
        active_set(root)
    elif which == "generic":
        generic(root)
        active_set(root)
        preconditioners(root)
        dense_joint(root)
        uamp_default(root)
        options(root)
    elif which == "headline":
        headline(root)
        generic(root)
        active_set(root)
        preconditioners(root)
        dense_joint(root)
        uamp_default(root)
        options(root)
    elif which == "long_cg":
        long_cg(root)
    elif which == "masked":
        masked_cg(root)
        masked_plans(root)
    elif which == "joint":
        joint_cg(root)
    elif which == "offgrid":
        offgrid(root)
    else:
        main(root)
        long_cg(root)
        masked_cg(root)
        masked_plans(root)
        joint_cg(root)
        offgrid(root)
        offgrid_1d(root)
        offgrid_wide(root)
        headline(root)
        generic(root)
        active_set(root)
        preconditioners(root)
        dense_joint(root)
        uamp_default(root)
        options(root)
