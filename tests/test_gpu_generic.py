"""Generic graph plans and presets beyond the two locked profiles, against
the reference's own output (tests/golden/generic_*.npz, written by
tests/golden/make_golden.py) and the float64 oracle.

* ``make_cuda_graph_sbl_plan(data, grid)`` with its default
  ``SBLConfig.fast_hybrid_uamp``: the truncation diagnostic
  (_active_truncation_delta_bound, solver.py:3927-3958) on the plan's own
  FFT passes, and ``hard_pruning_accepted`` against
  ``uamp_polish_max_truncation_delta`` (both outcomes are fixtures);
* ``SBLConfig.fast_uamp``: one reduced evaluation with diffuse tail loading,
  no pruning, re-initialisation or evidence gate (solver.py:178-195,
  3627-3633, 4087-4106), eager and graph-captured.

Bar: relative L2 1e-4 on x with the identical polished support, variance
and precision on the support and summed over the grid, noise, evidence
trace, iteration count; the certificate to 1e-4 relative.
"""

import numpy as np
import pytest
import torch

import headline_golden as hg
from oracle import fbme_oracle as fo

pytestmark = pytest.mark.gpu


def _out(res):
    out = dict(x=res.x.cpu().numpy(), posterior_variance=res.posterior_variance.cpu().numpy(),
               precision=res.precision.cpu().numpy(), noise_variance=float(res.noise_variance),
               effective_dof=float(res.effective_dof), evidence_trace=res.evidence_trace.cpu().numpy(),
               iterations=res.iterations)
    if res.truncation_delta_bound is not None:
        out["truncation_delta_bound"] = float(res.truncation_delta_bound)
        out["hard_pruning_accepted"] = bool(res.hard_pruning_accepted)
    return out


@pytest.mark.parametrize("case", hg.GENERIC_HYBRID)
def test_default_graph_plan_vs_reference(device, case):
    from torch_sbl import SBLConfig, make_cuda_graph_sbl_plan
    print("case", case, flush=True)
    g, scene, over = hg.load_generic(case)
    data = torch.from_numpy(scene.y).to(device)
    config = SBLConfig.fast_hybrid_uamp(device=device, **over) if over else None
    plan = make_cuda_graph_sbl_plan(data, scene.grid, config=config, warmup=1)
    res = plan.replay(synchronize=True)
    out = _out(res)
    assert "truncation_delta_bound" in out
    rel = hg.compare_generic(g, out)
    n, m = int(np.prod(scene.grid)), int(np.prod(scene.samples))
    orc = fo.fbme_plan(scene.y, scene.grid, config=fo.PlanConfig.fast_hybrid(n, m))
    print(case, "vs reference", rel, "bound", out["truncation_delta_bound"], float(g["truncation_delta_bound"]),
          "oracle", float(orc["truncation_delta_bound"]), flush=True)
    assert abs(out["truncation_delta_bound"] / float(orc["truncation_delta_bound"]) - 1) < 1e-4
    assert float(np.linalg.norm(out["x"].reshape(-1) - orc["x"]) / np.linalg.norm(orc["x"])) < 1e-4
    # a second replay on the same data reproduces the first bitwise
    x0 = res.x.clone()
    b0 = res.truncation_delta_bound.clone()
    plan.replay(data, synchronize=True)
    assert torch.equal(res.x, x0) and torch.equal(res.truncation_delta_bound, b0)


@pytest.mark.parametrize("case", hg.GENERIC_UAMP)
@pytest.mark.parametrize("graph", [False, True])
def test_fast_uamp_preset_vs_reference(device, case, graph):
    from torch_sbl import SBLConfig, fast_sbl, make_cuda_graph_sbl_plan
    print("case", case, "graph", graph, flush=True)
    g, scene, _ = hg.load_generic(case)
    data = torch.from_numpy(scene.y).to(device)
    cfg = SBLConfig.fast_uamp(device=device)
    if graph:
        res = make_cuda_graph_sbl_plan(data, scene.grid, config=cfg, warmup=1).replay(synchronize=True)
    else:
        res = fast_sbl(data, scene.grid, config=cfg)
    out = _out(res)
    assert res.truncation_delta_bound is None      # no pruning: no certificate
    rel = hg.compare_generic(g, out)
    print(case, "vs reference", rel, flush=True)


def test_batched_default_plan_equals_single(device):
    """A batch of three scenes in one default graph plan: each scene's
    reconstruction and certificate bitwise equal to its own plan's."""
    from torch_sbl import make_cuda_graph_sbl_plan
    from fbme_amd import scenes
    batch = [scenes.uniform_scene(1070 + i, (64, 64), (40, 40), k=15, snr_db=12.0) for i in range(3)]
    data = torch.from_numpy(np.stack([s.y for s in batch])).to(device)
    res = make_cuda_graph_sbl_plan(data, (64, 64), warmup=1).replay(synchronize=True)
    for i, s in enumerate(batch):
        one = make_cuda_graph_sbl_plan(torch.from_numpy(s.y).to(device), (64, 64), warmup=1).replay(synchronize=True)
        assert torch.equal(res.x[i], one.x)
        assert torch.equal(res.truncation_delta_bound[i], one.truncation_delta_bound)


def _general_cfg(device, mode, **kw):
    from torch_sbl import SBLConfig
    base = dict(mode=mode, device=str(device), max_iter=6, min_iter=1, tol=0.0, return_history=False)
    if mode == "cg":
        base.update(cg_fixed_iterations=12, hutchinson_probes=8, auto_probe_scaling=False)
    base.update(kw)
    return SBLConfig(**base)


@pytest.mark.parametrize("mode,grid,samples,batch", [("cg", (64, 64), (40, 40), 2),
                                                      ("cg", (256, 256), (160, 160), 1),
                                                      ("dense", (32, 32), (24, 24), 2)])
def test_general_solver_graph_plan_equals_eager(device, mode, grid, samples, batch):
    """mode='cg' with a fixed PCG budget and mode='dense' captured as one
    graph: every replay, on the example and on new data, equals the eager
    solve of the same data bitwise (same kernels in the same order)."""
    from fbme_amd import scenes
    from torch_sbl import CUDAGraphSBLPlan, sbl_2d
    print("general graph plan", mode, grid, samples, batch, flush=True)
    cfg = _general_cfg(device, mode)
    sets = []
    for seed in (1500, 1510):
        sc = [scenes.uniform_scene(seed + i, grid, samples, k=max(6, grid[0] * grid[1] // 400), snr_db=15.0)
              for i in range(batch)]
        sets.append(torch.from_numpy(np.stack([s.y for s in sc]).astype(np.complex64)).to(device))
    plan = CUDAGraphSBLPlan(sets[0], grid, config=cfg, warmup=1)
    for data in (sets[1], sets[0], sets[1]):
        res = plan.replay(data, synchronize=True)
        ref = sbl_2d(data, grid, config=cfg)
        assert res.mode == ref.mode == mode
        assert torch.equal(res.x, ref.x)
        assert torch.equal(res.posterior_variance, ref.posterior_variance)
        assert torch.equal(res.noise_variance, ref.noise_variance)
    assert not plan.factorization_failed()


def test_general_solver_graph_plan_refuses_host_reads(device):
    """The captures the reference cannot make either fail with RuntimeError:
    PCG with early exit (cg_fixed_iterations=0), the single-scene
    active-set switch within max_iter, convergence checks."""
    from fbme_amd import scenes
    from torch_sbl import CUDAGraphSBLPlan
    sc = scenes.uniform_scene(1520, (64, 64), (40, 40), k=10, snr_db=15.0)
    data = torch.from_numpy(sc.y[None].astype(np.complex64)).to(device)
    for kw in (dict(cg_fixed_iterations=0), dict(max_iter=40)):
        with pytest.raises(RuntimeError, match="Graph capture failed"):
            CUDAGraphSBLPlan(data, (64, 64), config=_general_cfg(device, "cg", **kw), warmup=1)
    with pytest.raises(ValueError):
        CUDAGraphSBLPlan(data, (64, 64), config=_general_cfg(device, "cg", tol=1e-4), warmup=1)
