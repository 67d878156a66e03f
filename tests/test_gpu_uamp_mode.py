"""sbl_2d(mode='uamp') with the SBLConfig defaults on the native plan: the
reference's convergence checks (tol 1e-4 from iteration 10, every 2
iterations, patience 2; solver.py:1285-1312, 1800-1904) driven from the host
over the plan's phases, and no polish (uamp_polish_size = 0: the UAMP
posterior of the final hyperparameters, solver.py:1907-1925).

Against the reference's runs (tests/golden/uamp_*.npz, written by
tests/golden/make_golden.py) and the float64 oracle: the same iteration
count and convergence decision, x and the variances to 1e-4 relative, noise
and dof to 1e-4, the per-iteration change and noise histories to 1e-3; the
same solve without a history (changes tracked at the checks only) is
bitwise the same, and a fixed-budget no-polish plan captured as a graph
replays bitwise like the eager solve.
"""

import numpy as np
import pytest
import torch
from conftest import golden_path

from oracle import fbme_oracle as fo
from test_oracle_golden import UAMP_MODE_CASES, uamp_mode_config

pytestmark = pytest.mark.gpu


def _config(device, g, **kw):
    from torch_sbl import SBLConfig
    over = {k[len("override_"):]: g[k].item() for k in g.files if k.startswith("override_")}
    over.update(kw)
    return SBLConfig(mode="uamp", device=str(device), **over)


@pytest.mark.parametrize("case", UAMP_MODE_CASES)
def test_uamp_mode_vs_reference_and_oracle(device, case):
    from torch_sbl import sbl_2d
    print("case", case, flush=True)
    g = np.load(golden_path(case + ".npz"))
    grid, samples, cfg = uamp_mode_config(g)
    y = torch.from_numpy(g["y"]).to(device)
    res = sbl_2d(y, grid, config=_config(device, g, return_history=True))
    b = y.shape[0]
    x = res.x.reshape(b, -1).cpu().numpy()
    ref = g["x"].reshape(b, -1)
    orc = fo.fbme_plan(g["y"].astype(np.complex128), grid, config=cfg)
    ox = np.asarray(orc["x"]).reshape(b, -1)
    rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    rel_o = np.linalg.norm(x - ox) / np.linalg.norm(ox)
    print(f"  iterations {res.iterations} (reference {int(g['iterations'])}), converged {res.converged}, "
          f"x vs reference {rel:.2e}, vs oracle {rel_o:.2e}", flush=True)
    assert res.iterations == int(g["iterations"])
    assert res.converged == bool(g["converged"])
    assert rel < 1e-4 and rel_o < 1e-4
    v = res.posterior_variance.reshape(b, -1).cpu().numpy()
    assert np.linalg.norm(v - g["posterior_variance"].reshape(b, -1)) / np.linalg.norm(g["posterior_variance"]) < 1e-4
    np.testing.assert_allclose(res.noise_variance.reshape(-1).cpu().numpy(), g["noise_variance"].reshape(-1),
                               rtol=1e-4)
    np.testing.assert_allclose(res.effective_dof.reshape(-1).cpu().numpy(), g["effective_dof"].reshape(-1),
                               rtol=1e-4)
    np.testing.assert_allclose(res.precision.reshape(b, -1).cpu().numpy(), g["precision"].reshape(b, -1),
                               rtol=1e-3)
    np.testing.assert_allclose(res.history["relative_change"], g["history_relative_change"], rtol=1e-3)
    np.testing.assert_allclose(res.history["noise_variance"], g["history_noise_variance"], rtol=1e-4)
    assert abs(res.relative_change / float(g["relative_change"]) - 1) < 1e-3
    if int(g["active_set_size"]) == 0:
        assert res.active_set_size == 0 and res.negative_log_evidence is None
    else:
        assert np.array_equal(np.flatnonzero(x), np.flatnonzero(ref))
    # without a history the changes are tracked at the checks only: same result
    quiet = sbl_2d(y, grid, config=_config(device, g, return_history=False))
    assert quiet.iterations == res.iterations and quiet.converged == res.converged
    assert torch.equal(quiet.x, res.x) and torch.equal(quiet.noise_variance, res.noise_variance)
    assert quiet.history == {}


def test_fixed_budget_no_polish_graph_plan(device):
    """tol = 0 and no polish: the plan runs as one call and captures into a
    graph; replays equal the eager solve bitwise."""
    from torch_sbl import CUDAGraphSBLPlan, SBLConfig, sbl_2d
    g = np.load(golden_path("uamp_64x64_40x40.npz"))
    y = torch.from_numpy(g["y"]).to(device)
    cfg = SBLConfig(mode="uamp", device=str(device), tol=0.0, return_history=False)
    plan = CUDAGraphSBLPlan(y, (64, 64), config=cfg, warmup=1)
    res = plan.replay(y, synchronize=True)
    ref = sbl_2d(y, (64, 64), config=cfg)
    assert res.iterations == ref.iterations == int(g["iterations"])
    assert torch.equal(res.x, ref.x) and torch.equal(res.posterior_variance, ref.posterior_variance)
    assert torch.equal(res.effective_dof, ref.effective_dof)
    rel = float(torch.linalg.norm(res.x.reshape(-1).cpu() - torch.from_numpy(g["x"].reshape(-1)))
                / np.linalg.norm(g["x"]))
    assert rel < 1e-4
