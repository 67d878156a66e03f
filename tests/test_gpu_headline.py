"""Headline workloads on the GPU against the reference's own output.

* The locked Fast / Stable plans at 256x256 / 160x160 (K=300 ellipse,
  10 dB, three seeds) and the Stable plan at 512x512 / 384x256 (K=1200
  clustered, 0 / 10 / 20 dB) against fixtures the reference's CPU Torch
  backend wrote (tests/golden/headline_*.npz; bar in tests/headline_golden.py:
  relative L2 <= 1e-4 on x, identical polished support, support mask,
  exchange and iteration counts).
* The 64-scene uneven batch of the benchmark (grids 64^2..256^2, ratios
  0.3-0.8, seed 2024) in one graph-captured plan: every scene bitwise equal
  to its single-scene plan, and two scenes of every grid size against the
  float64 oracle (relative L2 <= 1e-4, identical support and mask).
"""

import numpy as np
import pytest
import torch

import headline_golden
from fbme_amd import scenes
from oracle import fbme_oracle as fo

pytestmark = pytest.mark.gpu


def _plan(profile, data, grid):
    from torch_sbl_fastest_stable import make_fastest_plan, make_stable_plan
    fn = make_fastest_plan if profile == "fast" else make_stable_plan
    return fn(data, grid, polish_size=32, warmup=1)


@pytest.mark.parametrize("case,profile", headline_golden.CASES)
def test_headline_plan_vs_reference(device, case, profile):
    print("case", case, profile, flush=True)
    g, scene = headline_golden.load(case)
    data = torch.from_numpy(scene.y).to(device)
    plan = _plan(profile, data, scene.grid)
    res = plan.replay(synchronize=True)
    out = dict(x=res.x.cpu().numpy(), posterior_variance=res.posterior_variance.cpu().numpy(),
               support_mask=res.support_mask.cpu().numpy(),
               noise_variance=float(res.noise_variance), effective_dof=float(res.effective_dof),
               evidence_trace=res.evidence_trace.cpu().numpy(), iterations=res.iterations)
    if profile == "stable":
        out.update(support_exchanges_accepted=int(res.support_exchanges_accepted),
                   fixed_budget_gamma_residual=float(res.fixed_budget_gamma_residual),
                   fixed_budget_monotone=bool(res.fixed_budget_monotone))
    rel = headline_golden.compare(g, profile, out)
    print(case, profile, "GPU vs reference relative L2", rel, flush=True)
    assert int(plan.native_plan.out.status[0]) == 0


@pytest.mark.parametrize("profile", ["fast", "stable"])
def test_uneven_64_scene_batch(device, profile):
    from fbme_amd.batched import UnevenPlan
    from fbme_amd.plan import NativePlan
    batch = scenes.uneven_batch(2024, count=64)
    shapes = [(s.grid[0], s.grid[1], s.samples[0], s.samples[1]) for s in batch]
    print("uneven 64 scenes", profile, flush=True)
    plan = UnevenPlan(shapes, device=device, profile=profile)
    plan.replay([torch.from_numpy(s.y).to(device) for s in batch], synchronize=True)
    assert not plan.plan.out.status.any()
    per_grid = {}
    for i, s in enumerate(batch):
        got = plan.scene(i)
        single = NativePlan([shapes[i]], plan.options, device)
        o1 = single.run(torch.from_numpy(s.y.reshape(-1)).to(device))
        torch.cuda.synchronize()
        assert torch.equal(got["x"].reshape(-1), o1.x), f"scene {i} {shapes[i]}"
        assert torch.equal(got["posterior_variance"].reshape(-1), o1.posterior_variance)
        assert torch.equal(got["noise_variance"], o1.noise[0])
        assert torch.equal(got["support_mask"].reshape(-1), o1.support_mask)
        per_grid.setdefault(s.grid[0], []).append(i)
        del single
    assert sorted(per_grid) == [64, 128, 192, 256]
    checked = 0
    for g, idx in sorted(per_grid.items()):
        for i in idx[:2]:
            s = batch[i]
            got = plan.scene(i)
            x = got["x"].reshape(-1).cpu().numpy()
            orc = fo.fbme_plan(s.y, s.grid, profile)
            rel = float(np.linalg.norm(x - orc["x"]) / np.linalg.norm(orc["x"]))
            print("scene", i, shapes[i], "vs oracle", rel, flush=True)
            assert rel < 1e-4
            assert set(np.flatnonzero(x)) == set(np.flatnonzero(orc["x"]))
            assert np.array_equal(got["support_mask"].reshape(-1).cpu().numpy(), orc["support_mask"])
            assert abs(float(got["noise_variance"]) / float(orc["noise_variance"]) - 1) < 1e-4
            checked += 1
    assert checked == 8
