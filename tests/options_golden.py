"""Reference runs with SBLConfig options the presets leave at their defaults
(tests/golden/options_*.npz, written by tests/golden/make_golden.py):
a fixed noise variance (solver.py:1151-1156, 1747-1750, 1788-1798,
3457-3458, 4358-4387) under the Fast / Stable plan numerics and in mode
'cg' (also with joint sparsity), and uniform initial variances
(solver.py:1070-1089), randomised Hadamard probes in mode 'cg'
(solver.py:6002-6083) and cg_matvec='fourier' (Q d through full-grid FFTs,
operators.py:299-371: the same operator as the BTTB products the solver
runs), low-rank preconditioners of rank 8 and 0 (solver.py:5786-5848) and
the randomised Nystrom preconditioner (solver.py:5851-5976) at ranks 32 and
12, also with joint sparsity and a shared noise, adaptive probe counts
(16, 24, then all; solver.py:6086-6138) and the pattern-coupled prior with
a fixed coupling (solver.py:1100-1150, 1489-1590, 2396-2720)."""

import numpy as np

from conftest import golden_path
from oracle import cg_oracle as co
from oracle import fbme_oracle as fo

PLAN = ["options_stable_fixed_noise_64x64_40x40", "options_fast_fixed_noise_96x80_60x48",
        "options_fast_uniform_init_96x80_60x48"]
CG = ["options_cg_fixed_noise_32x32_24x24", "options_cg_hadamard_32x32_24x24", "options_cg_hadamard_48x40_40x30",
      "options_cg_fourier_32x32_24x24", "options_cg_rank8_48x40_40x30", "options_cg_rank0_32x32_24x24",
      "options_cg_nystrom_32x32_24x24", "options_cg_nystrom_rank12_64x64_40x40",
      "options_cg_adaptive_probes_64x64_40x40", "options_cg_spatial_zero_32x32_24x24",
      "options_cg_spatial_circular_48x40_40x30"]
CG_JOINT = ["options_cg_joint_fixed_noise_32x32_24x24_b3", "options_cg_joint_nystrom_32x32_24x24_b3",
            "options_cg_joint_adaptive_probes_32x32_24x24_b2"]


def load(name):
    g = np.load(golden_path(name + ".npz"))
    over = {k[len("override_"):]: g[k].item() for k in g.files if k.startswith("override_")}
    return g, tuple(int(v) for v in g["grid"]), over


def plan_config(g, over):
    """fo.PlanConfig of the golden's profile and options."""
    cfg = fo.PlanConfig.stable() if str(g["entry"]) == "stable" else fo.PlanConfig.fast()
    if "fixed_noise_variance" in over:
        cfg.fixed_noise_variance = float(over["fixed_noise_variance"])
    if over.get("initialization") == "uniform":
        cfg.initialization_blend = 0.0
    return cfg


def cg_config(over):
    kw = {k: v for k, v in over.items() if k in ("fixed_noise_variance", "max_iter", "probe_type",
                                                  "hutchinson_probes", "auto_probe_scaling",
                                                  "cg_preconditioner_rank", "cg_preconditioner",
                                                  "adaptive_probes", "spatial_coupling", "spatial_boundary")}
    return co.CGConfig(**kw)


def rel(a, b):
    a, b = np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def compare_plan(g, out):
    """x and the posterior variance to 1e-4 relative with the identical
    support and support mask, noise and dof, evidence trace, the Stable
    exchange count and fixed-budget residuals (zero noise residual with a
    fixed noise)."""
    assert rel(out["x"], g["x"]) < 1e-4
    assert set(np.flatnonzero(np.asarray(out["x"]).reshape(-1))) == set(np.flatnonzero(g["x"].reshape(-1)))
    assert np.array_equal(np.asarray(out["support_mask"]).reshape(-1), g["support_mask"].reshape(-1))
    assert rel(out["posterior_variance"], g["posterior_variance"]) < 1e-4
    assert abs(float(out["noise_variance"]) / float(g["noise_variance"]) - 1) < 1e-4
    assert abs(float(out["effective_dof"]) - float(g["effective_dof"])) < 1e-3
    assert np.allclose(np.asarray(out["evidence_trace"]).reshape(-1), g["evidence_trace"].reshape(-1),
                       rtol=2e-4, atol=0.1)
    if "support_exchanges_accepted" in g.files:
        assert int(out["support_exchanges_accepted"]) == int(g["support_exchanges_accepted"])
        assert abs(float(out["fixed_budget_gamma_residual"]) - float(g["fixed_budget_gamma_residual"])) < 1e-4
        assert float(out["fixed_budget_noise_residual"]) == float(g["fixed_budget_noise_residual"])
