"""SBLConfig options beyond the presets' defaults on the HIP path, against
the reference's runs (tests/options_golden.py) and the float64 oracles:

* fixed_noise_variance under the Fast and Stable plan numerics (the UAMP
  sweep starts at the fixed value and clamps it at the floor after every
  update, the polish keeps it, the Stable noise residual is zero), eagerly
  and as a graph plan (replays bitwise equal to the eager solve);
* initialization='uniform' under the Fast numerics;
* fixed_noise_variance in mode 'cg', also with joint sparsity and a shared
  noise, and randomised Hadamard probes (one at 36 active-set atoms): the
  same iteration, PCG and active-set counts as the reference, x to 1e-4.
"""

import numpy as np
import pytest
import torch

import options_golden as og
from oracle import cg_oracle as co
from oracle import fbme_oracle as fo

pytestmark = pytest.mark.gpu

PLAN_OVER = dict(
    fast=dict(uamp_polish_size=32, uamp_polish_iterations=2, uamp_polish_residual_exchange_size=0,
              uamp_polish_fixed_budget_diagnostics=False, uamp_polish_compute_truncation_diagnostic=False),
    stable=dict(uamp_polish_size=32, uamp_polish_iterations=3, uamp_polish_residual_exchange_size=12,
                uamp_polish_fixed_budget_diagnostics=True, uamp_polish_compute_truncation_diagnostic=False))


def _out(res):
    out = {name: getattr(res, name) for name in (
        "x", "posterior_variance", "support_mask", "noise_variance", "effective_dof", "evidence_trace",
        "support_exchanges_accepted", "fixed_budget_gamma_residual", "fixed_budget_noise_residual")}
    return {k: (v.cpu().numpy() if torch.is_tensor(v) else v) for k, v in out.items() if v is not None}


@pytest.mark.parametrize("case", og.PLAN)
def test_plan_options_vs_reference_and_oracle(device, case):
    from torch_sbl import SBLConfig, fast_sbl, make_cuda_graph_sbl_plan
    print("case", case, flush=True)
    g, grid, over = og.load(case)
    y = torch.from_numpy(g["y"][0]).to(device)
    cfg = SBLConfig.paper_hybrid_uamp(device=str(device), **PLAN_OVER[str(g["entry"])], **over)
    res = fast_sbl(y, grid, config=cfg)
    out = _out(res)
    og.compare_plan(g, out)
    orc = fo.fbme_plan(g["y"][0].astype(np.complex128), grid, config=og.plan_config(g, over))
    print(case, "x vs reference", og.rel(out["x"], g["x"]), "vs oracle", og.rel(out["x"], orc["x"]), flush=True)
    assert og.rel(out["x"], orc["x"]) < 1e-4
    if "fixed_noise_variance" in over:
        assert float(out["noise_variance"]) == float(np.float32(over["fixed_noise_variance"]))
    # the same options in a graph plan: replays equal the eager solve
    plan = make_cuda_graph_sbl_plan(y, grid, config=cfg, warmup=1)
    rep = plan.replay(y, synchronize=True)
    assert torch.equal(rep.x, res.x) and torch.equal(rep.posterior_variance, res.posterior_variance)
    assert torch.equal(rep.noise_variance, res.noise_variance)


def _cg(device, g, grid, over, joint=False):
    from torch_sbl import SBLConfig, sbl_2d
    y = torch.from_numpy(g["y"] if joint else g["y"][0]).to(device)
    cfg = SBLConfig(mode="cg", device=str(device), **over)
    return sbl_2d(y, grid, config=cfg)


@pytest.mark.parametrize("case", og.CG + og.CG_JOINT)
def test_cg_fixed_noise_vs_reference(device, case):
    print("case", case, flush=True)
    g, grid, over = og.load(case)
    joint = case in og.CG_JOINT
    res = _cg(device, g, grid, over, joint)
    x = res.x.cpu().numpy()
    rel = og.rel(x, g["x"])
    print(case, "iterations", res.iterations, int(g["iterations"]), "x vs reference", rel, flush=True)
    assert res.iterations == int(g["iterations"])
    assert res.active_set_size == int(g["active_set_size"])
    assert rel < 1e-4
    np.testing.assert_array_equal(np.asarray(res.history["cg_iterations"]), g["history_cg_iterations"])
    noise = res.noise_variance.reshape(-1).cpu().numpy()
    if "fixed_noise_variance" in over:
        np.testing.assert_allclose(noise, g["noise_variance"].reshape(-1), rtol=1e-6)
        assert np.all(noise == np.float32(over["fixed_noise_variance"]))
        np.testing.assert_allclose(res.history["noise_variance"], g["history_noise_variance"], rtol=1e-6)
    else:
        np.testing.assert_allclose(noise, g["noise_variance"].reshape(-1), rtol=1e-4)
        np.testing.assert_allclose(res.history["noise_variance"], g["history_noise_variance"], rtol=1e-4)
    if not joint:
        orc = co.sbl_cg(g["y"][0], grid, og.cg_config(over))
        assert og.rel(x, orc["x"]) < 1e-3
