"""The float64 oracle against golden vectors written by the reference's CPU
Torch backend (tests/golden/make_golden.py).

Tolerances: the reference computes in complex64 with float32 atom phases,
so x agrees with the exact float64 oracle to ~1e-5 relative L2 on these
scenes; the test uses 1e-4 (the hot path's stated parity bar) and requires
identical supports, masks and exchange decisions.
"""

import numpy as np
import pytest

from conftest import golden_path
from oracle import fbme_oracle as fo

import headline_golden

CASES = ["plan_32x32_24x24", "plan_40x32_20x16", "plan_64x64_40x40",
         "plan_96x80_60x48",
         # observation masks (fast_sbl(..., observation_mask=), 70% / 60%)
         "masked_plan_64x64_40x40", "masked_plan_96x80_60x48"]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("profile", ["fast", "stable"])
def test_oracle_matches_reference_golden(case, profile):
    g = np.load(golden_path(case + ".npz"))
    grid = tuple(int(v) for v in g["grid"])
    out = fo.fbme_plan(g["y"], grid, profile, mask=g["mask"] if "mask" in g else None)
    xr = g[profile + "_x"].reshape(-1)
    assert _rel(out["x"], xr) < 1e-4
    assert set(np.flatnonzero(out["x"])) == set(np.flatnonzero(xr))
    assert np.array_equal(out["support_mask"],
                          g[profile + "_support_mask"].reshape(-1))
    assert abs(out["noise_variance"] / g[profile + "_noise_variance"] - 1) < 1e-4
    assert _rel(out["posterior_variance"],
                g[profile + "_posterior_variance"].reshape(-1)) < 1e-4
    assert abs(out["effective_dof"] - g[profile + "_effective_dof"]) < 1e-3
    trace = g[profile + "_evidence_trace"].reshape(-1)
    assert np.allclose(out["evidence_trace"], trace, rtol=2e-4, atol=0.1)
    assert int(out["iterations"]) == int(g[profile + "_iterations"])
    if profile == "stable":
        assert int(out["support_exchanges_accepted"]) == int(
            g["stable_support_exchanges_accepted"])
        assert abs(out["fixed_budget_gamma_residual"]
                   - g["stable_fixed_budget_gamma_residual"]) < 1e-4
        assert abs(out["fixed_budget_noise_residual"]
                   - g["stable_fixed_budget_noise_residual"]) < 1e-4
        assert bool(out["fixed_budget_monotone"]) == bool(
            g["stable_fixed_budget_monotone"])


def test_planted_mistake_is_caught():
    """A plausible slip (UAMP damping 0.1 instead of 0.05, or a dropped
    noise*dof term) must fail the golden comparison."""
    g = np.load(golden_path("plan_64x64_40x40.npz"))
    grid = tuple(int(v) for v in g["grid"])
    y = g["y"].reshape(1, -1).astype(np.complex128)
    samples = tuple(int(v) for v in g["samples"])
    cfg = fo.PlanConfig.fast()
    cfg.uamp_damping = 0.1
    sweep = fo.uamp_sweep(y, grid, samples, cfg)
    res = fo.polish(y, sweep, grid, samples, cfg)
    assert _rel(res["mean"][0], g["fast_x"].reshape(-1)) > 1e-3
    cfg = fo.PlanConfig.fast()
    cfg.noise_update = "evidence"
    sweep = fo.uamp_sweep(y, grid, samples, cfg)
    res = fo.polish(y, sweep, grid, samples, cfg)
    assert _rel(res["mean"][0], g["fast_x"].reshape(-1)) > 1e-4


def test_full_mask_is_unmasked_plan():
    g = np.load(golden_path("plan_40x32_20x16.npz"))
    grid = tuple(int(v) for v in g["grid"])
    a = fo.fbme_plan(g["y"], grid, "stable")
    b = fo.fbme_plan(g["y"], grid, "stable", mask=np.ones(g["y"].shape, bool))
    np.testing.assert_allclose(b["x"], a["x"], rtol=1e-12, atol=1e-14)
    assert np.array_equal(b["support"], a["support"])


def test_masked_gram_apply_is_dense():
    """The masked Gram pieces against the dense row-restricted matrix."""
    grid, samples = (12, 10), (7, 6)
    mask = np.random.default_rng(8).random(samples) < 0.6
    a = fo.dense_prefix_matrix(grid, samples)[mask.reshape(-1)]
    sup = np.array([[3, 17, 40, 99]])
    np.testing.assert_allclose(fo.masked_gram(sup, grid, samples, mask)[0],
                               a[:, sup[0]].conj().T @ a[:, sup[0]], atol=1e-12)
    mu = np.array([[1 + 2j, -0.5j, 0.25, 3.0]])
    np.testing.assert_allclose(fo._gram_apply(mu, sup, grid, samples, mask)[0],
                               a.conj().T @ (a[:, sup[0]] @ mu[0]), atol=1e-11)


@pytest.mark.parametrize("case,profile", headline_golden.CASES)
def test_oracle_matches_reference_at_headline_sizes(case, profile):
    """256x256 / 160x160 (K=300 ellipse, 10 dB, three seeds) and 512x512 /
    384x256 (K=1200 clustered, 0/10/20 dB): K >> polish_size, where top-k
    boundaries, evidence gates and exchanges meet many strong off-support
    atoms and the reference's float32 atom phases are least accurate."""
    g, scene = headline_golden.load(case)
    out = fo.fbme_plan(scene.y, scene.grid, profile)
    rel = headline_golden.compare(g, profile, out, trace_rtol=2e-4, trace_atol=0.1)
    print(case, profile, "oracle vs reference", rel)


def _generic_config(g, scene, overrides):
    n = int(np.prod(scene.grid))
    m = int(np.prod(scene.samples))
    preset = str(g["preset"])
    return (fo.PlanConfig.fast_hybrid if preset == "fast_hybrid_uamp" else fo.PlanConfig.fast_uamp)(n, m)


@pytest.mark.parametrize("case", headline_golden.GENERIC_HYBRID + headline_golden.GENERIC_UAMP)
def test_oracle_matches_reference_generic_presets(case):
    """SBLConfig.fast_hybrid_uamp (make_cuda_graph_sbl_plan's default: four
    UAMP steps at 256x256 / 160x160, evidence margin 0, hyperpriors 1e-8,
    truncation diagnostic) and SBLConfig.fast_uamp (13 UAMP steps at
    256x256, one reduced evaluation with diffuse tail loading and no
    pruning / re-initialisation / gate) against the reference."""
    g, scene, overrides = headline_golden.load_generic(case)
    cfg = _generic_config(g, scene, overrides)
    out = fo.fbme_plan(scene.y, scene.grid, config=cfg)
    delta = overrides.get("uamp_polish_max_truncation_delta")
    if delta is not None:
        out["hard_pruning_accepted"] = out["truncation_delta_bound"] <= delta
    rel = headline_golden.compare_generic(g, out)
    print(case, "oracle vs reference", rel)


def test_truncation_bound_is_the_dense_tail_covariance_norm():
    """The certificate's numerator equals the largest eigenvalue magnitude
    of the circulant embedding, which bounds the spectral norm of the dense
    tail covariance A Gamma_tail A^H (a principal submatrix of it): checked
    against the dense matrix on a small grid, and the embedding against its
    definition (a zero tail gives a zero bound)."""
    grid, samples = (12, 10), (7, 6)
    rng = np.random.default_rng(3)
    gamma = rng.exponential(size=(1, 120))
    sup = np.array([[4, 50, 77]])
    noise = np.array([0.5])
    bound = fo.truncation_delta_bound(gamma, noise, sup, grid, samples)[0]
    a = fo.dense_prefix_matrix(grid, samples)
    tail = gamma[0].copy()
    tail[sup[0]] = 0.0
    cov = (a * tail) @ a.conj().T
    norm = np.linalg.norm(cov, 2)
    assert norm <= bound * noise[0] * (1 + 1e-12)
    assert bound * noise[0] <= 2.0 * norm          # the embedding is not loose here
    assert fo.truncation_delta_bound(np.zeros((1, 120)), noise, sup, grid, samples)[0] == 0.0


@pytest.mark.parametrize("case", ["plan_64x64_40x40", "plan_96x80_60x48", "headline_256x256_160x160_s1001"])
def test_torch_cpu_baseline_port_matches_oracle(case):
    """The benchmark's CPU baseline (oracle/torch_cpu_plan.py: the Fast plan
    on torch's CPU kernels, complex64 sweep) against the float64 oracle."""
    import torch
    from oracle import torch_cpu_plan as tcp
    if case.startswith("headline"):
        _, scene = headline_golden.load(case)
        y, grid = scene.y, scene.grid
    else:
        g = np.load(golden_path(case + ".npz"))
        y, grid = g["y"], tuple(int(v) for v in g["grid"])
    out = tcp.fast_plan(torch.from_numpy(np.stack([y, y]))[:1], grid)
    ref = fo.fbme_plan(y, grid, "fast")
    x = out["x"].numpy().reshape(-1)
    rel = _rel(x, ref["x"])
    print(case, "torch CPU port vs oracle", rel)
    assert rel < 1e-4
    assert set(np.flatnonzero(x)) == set(np.flatnonzero(ref["x"]))
    assert abs(float(out["noise"][0]) / float(ref["noise_variance"]) - 1) < 1e-4


UAMP_MODE_CASES = ["uamp_64x64_40x40", "uamp_128x96_48x40_b2", "uamp_64x48_60x44_tol1e-3",
                   "uamp_64x64_40x40_polish", "uamp_64x64_40x40_uniform", "uamp_96x64_48x40_fixed_budget"]
_UAMP_FIELDS = dict(tol="tol", max_iter="max_iter", uamp_iteration_offset="uamp_iteration_offset",
                    uamp_polish_size="polish_size", uamp_auto_iterations="uamp_auto_iterations")


def uamp_mode_config(g):
    """PlanConfig of SBLConfig(mode='uamp') with the golden's overrides
    (initialization='uniform' is the backprojection blend 0,
    solver.py:1070-1089)."""
    grid, samples = tuple(int(v) for v in g["grid"]), tuple(int(v) for v in g["samples"])
    over = {}
    for key in g.files:
        if not key.startswith("override_"):
            continue
        name, value = key[len("override_"):], g[key].item()
        if name == "initialization":
            over["initialization_blend"] = 0.1 if value == "backprojection" else 0.0
        else:
            over[_UAMP_FIELDS[name]] = value
    return grid, samples, fo.PlanConfig.uamp_default(int(np.prod(grid)), int(np.prod(samples)), **over)


@pytest.mark.parametrize("case", UAMP_MODE_CASES)
def test_oracle_matches_reference_uamp_mode(case):
    """sbl_2d(mode='uamp') with the SBLConfig defaults (convergence checks at
    tol, no polish) and with a 16-atom polish: same iteration count and
    convergence decision, x and variances to 1e-4 relative (the polish case:
    identical support), noise and dof to 1e-4."""
    g = np.load(golden_path(case + ".npz"))
    grid, samples, cfg = uamp_mode_config(g)
    out = fo.fbme_plan(g["y"].astype(np.complex128), grid, config=cfg)
    assert int(out["iterations"]) == int(g["iterations"])
    b = g["y"].shape[0]
    x = np.asarray(out["x"]).reshape(b, -1)
    ref = g["x"].reshape(b, -1)
    rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    print(case, "iterations", int(g["iterations"]), "x rel", rel)
    assert rel < 1e-4
    v = np.asarray(out["posterior_variance"]).reshape(b, -1)
    assert np.linalg.norm(v - g["posterior_variance"].reshape(b, -1)) / np.linalg.norm(g["posterior_variance"]) < 1e-4
    np.testing.assert_allclose(np.atleast_1d(out["noise_variance"]), g["noise_variance"].reshape(-1), rtol=1e-4)
    np.testing.assert_allclose(np.atleast_1d(out["effective_dof"]), g["effective_dof"].reshape(-1), rtol=1e-4)
    if cfg.polish_size == 0:
        assert bool(out["converged"]) == bool(g["converged"])
        assert abs(out["relative_change"] / float(g["relative_change"]) - 1) < 1e-3
    else:
        assert np.array_equal(np.flatnonzero(x), np.flatnonzero(ref))
